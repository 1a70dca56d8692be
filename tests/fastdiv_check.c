/* Checks the traversal kernel's tile division against IEEE division.
 *
 * The kernel computes p / size, size = d * 2^k (d = 1..255, k in {0,3,7,12}),
 * as q0 = RN(p*y), r = fma(-q0, size, p), q = fma(r, y, q0) with y = RN(1/d)*2^-k,
 * and uses it only when 2^-64 <= |p| <= 2^100. This program compares q and the
 * modulo p - size*floor(q) bit for bit with IEEE p/size for every divisor and
 * every stride-th significand of p at each biased exponent in [e0, e1].
 * Usage: fastdiv_check e0 e1 stride   -> prints "mismatches N" (exit 1 if N > 0)
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static float f_of(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t u_of(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

int main(int argc, char **argv) {
    if (argc != 4) return 2;
    const int e0 = atoi(argv[1]), e1 = atoi(argv[2]), stride = atoi(argv[3]);
    const int ks[4] = {0, 3, 7, 12};
    long tests = 0, bad = 0;
    for (int ki = 0; ki < 4; ++ki)
        for (int d = 1; d <= 255; ++d) {
            volatile float dd = (float)d;
            const float size = ldexpf(dd, ks[ki]);
            const float y = ldexpf(1.0f / dd, -ks[ki]);
            for (int e = e0; e <= e1; ++e)
                for (uint32_t m = 0; m < (1u << 23); m += (uint32_t)stride)
                    for (int sgn = 0; sgn < 2; ++sgn) {
                        float p = f_of(((uint32_t)e << 23) | m);
                        if (sgn) p = -p;
                        const float a = fabsf(p);
                        if (!(a >= 0x1p-64f && a <= 0x1p100f)) continue;
                        const float qt = p / size;
                        const float q0 = p * y;
                        const float r = fmaf(-q0, size, p);
                        const float q = fmaf(r, y, q0);
                        ++tests;
                        if (u_of(q) != u_of(qt) || u_of(p - size * floorf(q)) != u_of(p - size * floorf(qt))) {
                            if (bad < 5) printf("d=%d k=%d p=%a ieee=%a fast=%a\n", d, ks[ki], p, qt, q);
                            ++bad;
                        }
                    }
        }
    printf("tests %ld mismatches %ld\n", tests, bad);
    return bad ? 1 : 0;
}
