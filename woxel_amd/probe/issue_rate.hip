// VALU issue cost on the GPU it runs on: cycles per wave64 instruction per SIMD for
// single opcodes and for alternating pairs. Every lane runs 16 independent chains of
// one instruction (or of two alternating ones) for K trips; 8,192 workgroups of 256
// lanes keep 32 waves on each CU. The figure assumes 1,024 SIMDs at 2.4 GHz (MI355X).
//
//   hipcc --offload-arch=gfx950 -Xclang -target-feature -Xclang -packed-fp32-ops -O3 \
//       -o /tmp/issue_rate woxel_amd/probe/issue_rate.hip && /tmp/issue_rate
//
// Used to check the VALU issue model of the traversal's roofline (4 cycles per wave64
// instruction); not part of the library.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

namespace {

constexpr int kBlocks = 256 * 32, kTrips = 4096;

// 16 chains of INSN (operands: %0 the chain, %1 a constant source).
#define ONE(NAME, INSN, T)                                                                    \
    __global__ __launch_bounds__(256) void NAME(unsigned *out, int trips, unsigned seed) {    \
        T a[16];                                                                              \
        for (int i = 0; i < 16; ++i) a[i] = (T)(seed + threadIdx.x + i);                      \
        const T b = (T)(seed * 3u);                                                           \
        for (int k = 0; k < trips; ++k) {                                                     \
            _Pragma("unroll") for (int i = 0; i < 16; ++i) asm volatile(INSN : "+v"(a[i]) : "v"(b)); \
        }                                                                                     \
        unsigned s = 0;                                                                       \
        for (int i = 0; i < 16; ++i) s += (unsigned)a[i];                                     \
        out[blockIdx.x * 256 + threadIdx.x] = s;                                              \
    }

// 16 chains, INSN1 and INSN2 alternating.
#define TWO(NAME, INSN1, INSN2, T)                                                            \
    __global__ __launch_bounds__(256) void NAME(unsigned *out, int trips, unsigned seed) {    \
        T a[16];                                                                              \
        for (int i = 0; i < 16; ++i) a[i] = (T)(seed + threadIdx.x + i);                      \
        const T b = (T)(seed * 3u);                                                           \
        for (int k = 0; k < trips; ++k) {                                                     \
            _Pragma("unroll") for (int i = 0; i < 16; i += 2) {                               \
                asm volatile(INSN1 : "+v"(a[i]) : "v"(b));                                    \
                asm volatile(INSN2 : "+v"(a[i + 1]) : "v"(b));                                \
            }                                                                                 \
        }                                                                                     \
        unsigned s = 0;                                                                       \
        for (int i = 0; i < 16; ++i) s += (unsigned)a[i];                                     \
        out[blockIdx.x * 256 + threadIdx.x] = s;                                              \
    }

ONE(k_add_f32, "v_add_f32 %0, %0, %1", float)
ONE(k_mul_f32, "v_mul_f32 %0, %0, %1", float)
ONE(k_fma_f32, "v_fma_f32 %0, %0, %1, %0", float)
ONE(k_fmac_f32, "v_fmac_f32 %0, %1, %1", float)
ONE(k_min_f32, "v_min_f32 %0, %0, %1", float)
ONE(k_max3_f32, "v_max3_f32 %0, %0, %1, %1", float)
ONE(k_floor_f32, "v_floor_f32 %0, %0", float)
ONE(k_cvt_flr, "v_cvt_flr_i32_f32 %0, %0", unsigned)
ONE(k_rcp_f32, "v_rcp_f32 %0, %0", float)
ONE(k_mov_b32, "v_mov_b32 %0, %1", unsigned)
ONE(k_add_u32, "v_add_u32 %0, %0, %1", unsigned)
ONE(k_and_b32, "v_and_b32 %0, %0, %1", unsigned)
ONE(k_xor_b32, "v_xor_b32 %0, %0, %1", unsigned)
ONE(k_lshlrev_b32, "v_lshlrev_b32 %0, %1, %0", unsigned)
ONE(k_lshrrev_b32, "v_lshrrev_b32 %0, %1, %0", unsigned)
ONE(k_bfe_u32, "v_bfe_u32 %0, %0, %1, 3", unsigned)
ONE(k_lshl_or_b32, "v_lshl_or_b32 %0, %0, 3, %1", unsigned)
ONE(k_bcnt_u32, "v_bcnt_u32_b32 %0, %0, %1", unsigned)
ONE(k_ffbh_u32, "v_ffbh_u32 %0, %0", unsigned)
ONE(k_mul_lo_u32, "v_mul_lo_u32 %0, %0, %1", unsigned)
TWO(k_add_mul_f32, "v_add_f32 %0, %0, %1", "v_mul_f32 %0, %0, %1", float)
TWO(k_add_min_f32, "v_add_f32 %0, %0, %1", "v_min_f32 %0, %0, %1", float)
TWO(k_add_xor_u32, "v_add_u32 %0, %0, %1", "v_xor_b32 %0, %0, %1", unsigned)
TWO(k_add_bfe_u32, "v_add_u32 %0, %0, %1", "v_bfe_u32 %0, %0, %1, 3", unsigned)

// 16 compares into 16 SGPR pairs (no chain: every compare reads the same inputs).
__global__ __launch_bounds__(256) void k_cmp_eq_u32(unsigned *out, int trips, unsigned seed) {
    const unsigned a = seed + threadIdx.x, b = seed * 3u;
    uint64_t r[16];
    for (int k = 0; k < trips; ++k) {
#pragma unroll
        for (int i = 0; i < 16; ++i) asm volatile("v_cmp_eq_u32_e64 %0, %1, %2" : "=s"(r[i]) : "v"(a), "v"(b));
    }
    uint64_t s = 0;
    for (int i = 0; i < 16; ++i) s ^= r[i];
    out[blockIdx.x * 256 + threadIdx.x] = (unsigned)s;
}

// 16 chains of selects under one lane mask.
__global__ __launch_bounds__(256) void k_cndmask_b32(unsigned *out, int trips, unsigned seed) {
    unsigned a[16];
    for (int i = 0; i < 16; ++i) a[i] = seed + threadIdx.x + i;
    const unsigned b = seed * 3u;
    const uint64_t m = __ballot(threadIdx.x & 1u);
    for (int k = 0; k < trips; ++k) {
#pragma unroll
        for (int i = 0; i < 16; ++i) asm volatile("v_cndmask_b32_e64 %0, %0, %1, %2" : "+v"(a[i]) : "v"(b), "s"(m));
    }
    unsigned s = 0;
    for (int i = 0; i < 16; ++i) s += a[i];
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

}  // namespace

int main() {
    unsigned *out = nullptr;
    if (hipMalloc(&out, (size_t)kBlocks * 256 * sizeof(unsigned)) != hipSuccess) return 1;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
    struct Case {
        const char *name;
        void (*fn)(unsigned *, int, unsigned);
    };
    const Case cases[] = {
        {"v_add_f32", k_add_f32}, {"v_mul_f32", k_mul_f32}, {"v_fma_f32", k_fma_f32}, {"v_fmac_f32", k_fmac_f32},
        {"v_min_f32", k_min_f32}, {"v_max3_f32", k_max3_f32}, {"v_floor_f32", k_floor_f32},
        {"v_cvt_flr_i32_f32", k_cvt_flr}, {"v_rcp_f32", k_rcp_f32}, {"v_mov_b32", k_mov_b32},
        {"v_add_u32", k_add_u32}, {"v_and_b32", k_and_b32}, {"v_xor_b32", k_xor_b32},
        {"v_lshlrev_b32", k_lshlrev_b32}, {"v_lshrrev_b32", k_lshrrev_b32}, {"v_bfe_u32", k_bfe_u32},
        {"v_lshl_or_b32", k_lshl_or_b32}, {"v_bcnt_u32_b32", k_bcnt_u32}, {"v_ffbh_u32", k_ffbh_u32},
        {"v_mul_lo_u32", k_mul_lo_u32}, {"v_cmp_eq_u32_e64", k_cmp_eq_u32}, {"v_cndmask_b32_e64", k_cndmask_b32},
        {"v_add_f32 | v_mul_f32", k_add_mul_f32}, {"v_add_f32 | v_min_f32", k_add_min_f32},
        {"v_add_u32 | v_xor_b32", k_add_xor_u32}, {"v_add_u32 | v_bfe_u32", k_add_bfe_u32},
    };
    const double insts = (double)kBlocks * 4.0 * kTrips * 16.0;
    for (const Case &c : cases) {
        float best = 1e30f;
        for (int rep = 0; rep < 3; ++rep) {
            if (hipEventRecord(e0, 0) != hipSuccess) return 2;
            hipLaunchKernelGGL(c.fn, dim3(kBlocks), dim3(256), 0, 0, out, kTrips, 3u);
            float ms = 0;
            if (hipEventRecord(e1, 0) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
                return 2;
            if (rep && ms < best) best = ms;  // the first launch warms up
        }
        std::printf("%-24s %7.3f ms  %5.2f cycles per wave64 instruction per SIMD\n", c.name, best,
                    best * 1e-3 * 2.4e9 * 1024 / insts);
    }
    return hipGetLastError() == hipSuccess ? 0 : 3;
}
