// Host-code sanitizer harness (built by tests/test_host_sanitizers.py with
// AddressSanitizer and UBSan): the .vdb reader + finalize on corrupted copies of an
// asset -- truncations, bit flips, overwritten header fields, random byte runs. Every
// corrupted file must be read or rejected with a status code, never crash.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>
#include "woxel_hip.h"

static std::vector<unsigned char> slurp(const char *p) {
    FILE *f = std::fopen(p, "rb");
    std::vector<unsigned char> d;
    if (!f) return d;
    std::fseek(f, 0, SEEK_END);
    long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    d.resize((size_t)n);
    if (std::fread(d.data(), 1, d.size(), f) != d.size()) d.clear();
    std::fclose(f);
    return d;
}
static void spit(const char *p, const std::vector<unsigned char> &d) {
    FILE *f = std::fopen(p, "wb");
    std::fwrite(d.data(), 1, d.size(), f);
    std::fclose(f);
}
int main(int argc, char **argv) {
    const int iters = argc > 2 ? std::atoi(argv[2]) : 200;
    std::vector<unsigned char> base = slurp(argv[1]);
    std::mt19937_64 rng(12345);
    int ok = 0, fail = 0;
    const char *tmp = argc > 3 ? argv[3] : "/tmp/fuzz_vdb.vdb";
    for (int it = 0; it < iters; ++it) {
        std::vector<unsigned char> d = base;
        const int mode = it % 4;
        if (mode == 0) d.resize(rng() % d.size());                      // truncation
        else if (mode == 1) { for (int k = 0; k < 8; ++k) d[rng() % d.size()] ^= (unsigned char)(1u << (rng() % 8)); }
        else if (mode == 2) {                                            // corrupt a 4/8-byte field in the header area
            size_t at = rng() % std::min<size_t>(d.size(), 4096);
            for (int k = 0; k < 8 && at + k < d.size(); ++k) d[at + k] = (unsigned char)rng();
        } else {                                                         // random bytes in the middle
            size_t at = rng() % d.size();
            for (int k = 0; k < 64 && at + k < d.size(); ++k) d[at + k] = (unsigned char)rng();
        }
        spit(tmp, d);
        wx_vdb *v = nullptr;
        wx_vdb_create(&v);
        wx_vdb_file_info info;
        int rc = wx_vdb_read_file(v, tmp, nullptr, &info);
        if (rc == WX_OK) {
            rc = wx_vdb_finalize(v, 2);
            ok += rc == WX_OK;
        } else {
            ++fail;
        }
        wx_vdb_destroy(v);
    }
    std::printf("%s: %d read+finalized, %d rejected\n", argv[1], ok, fail);
    return 0;
}
