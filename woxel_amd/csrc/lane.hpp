// Per-lane HDDA + grid-SDF traversal shared by the traversal and render kernels.
//
// The per-ray arithmetic restates hdda_ray (src/shaders/raycast.comp.wgsl:84-126)
// operation for operation, in FP32 with contraction off (-ffp-contract=off) and
// IEEE division and square root, so results are bit-identical to the scalar
// oracle. The lookups restate get_vdb_leaf_from_{leaf,nothing,node5,node4,node3}
// (raycast.comp.wgsl:360-494) over the packed device layout: one 16-byte record
// per internal-node lookup (child/value bits, popcount base, tile SDF) and one
// 4-bit slot per voxel lookup (0 = active, else SDF), so every tree level costs a
// single gather.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/woxel_hip.h"

namespace wxk {

constexpr uint32_t kMaxSteps = 1000u;  // HDDA_MAX_RAY_STEPS (raycast.comp.wgsl:82)
constexpr uint32_t kRun = 0xFFu;       // lane state: still stepping

struct GridDev {
    const int4 *root;     // n5 x {kx, ky, kz, index}
    const uint4 *n5;      // n5 x 4096 records
    const uint4 *n4;      // n4 x 256 records
    const uint32_t *leaf; // n3 x 64 words
    uint32_t n5count;
    float bound;
    uint32_t wide;        // 1: the leaf pool exceeds 4 GiB (more than 2^24 leaves)
    uint32_t pad;
    int4 root0;           // root[0]: a grid with one node-5 looks its key up in scalar registers
};
static_assert(sizeof(GridDev) == 64, "GridDev layout");

// Root tables of up to kRootLds entries are scanned from LDS by the kernels that
// trace one grid (a copy made at workgroup start); a grid with one node-5 needs no
// table at all (GridDev::root0), a larger table is scanned through the L1 cache.
constexpr uint32_t kRootLds = 16u;
// Stages the root table of g in s_root when it has 2..kRootLds entries; returns
// whether it did (the lookups then scan s_root, else root0 or g.root). Call
// before the workgroup's first barrier (init_tab / init_rcp), which publishes the
// copy. n5count is uniform.
__device__ __forceinline__ bool stage_root(const GridDev &g, int4 *s_root) {
    if (g.n5count <= 1u || g.n5count > kRootLds) return false;
    for (uint32_t i = threadIdx.x; i < g.n5count; i += blockDim.x) s_root[i] = g.root[i];
    return true;
}

__device__ __forceinline__ float smin(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float smax(float a, float b) { return a < b ? b : a; }

// f32 -> i32: truncate, saturate, NaN -> 0. This is exactly the rule of the
// hardware conversion, so it is one instruction (a C++ cast of an out-of-range
// value would be undefined and lets the compiler assume anything).
__device__ __forceinline__ int32_t f2i(float x) {
    int32_t r;
    asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// f2i(floorf(x)) in one instruction. The fused conversion equals floor + convert on
// every f32 input except the NaNs (checked over all 2^32 bit patterns on gfx950), so
// it is used only where the coordinate cannot be NaN.
__device__ __forceinline__ int32_t f2i_floor(float x) {
    int32_t r;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// One ray in flight. The lookup cache of the reference (VdbLeaf.parents) is kept
// as the last looked-up voxel: after every lookup, the cached node at each level
// below num_parents is the node containing that voxel. The ray origin is not
// kept: camera rays start at the (uniform) eye, explicit rays re-read it at the
// end, shadow rays keep it in LDS. The axis mask of the last step is kept as the
// nudge each axis received (+-4e-4 on the stepped axes, +-0 on the others).
struct RayLane {
    float dx, dy, dz, ix, iy, iz, px, py, pz;
    float ax, ay, az;      // last nudge per axis
    uint32_t s01, s23;     // steps[k] = lookups resolved at num_parents k: s01 = steps0 | steps1 << 16,
                           // s23 = steps2 | steps3 << 16
    uint32_t it, lk;       // iterations done (kept by lane_step<.., kScalarIt = false> only); lookup
                           // counters L5 | L4 << 10 (L3 = steps3: every leaf lookup ends the chain)
    uint32_t np, state;    // num_parents of the last lookup; kRun or the final state
    int32_t qx, qy, qz;    // last looked-up voxel
    uint32_t i5, i4, i3;   // cached node indices
};

// The hit record's lookup word L5 | L4 << 10 | L3 << 20.
__device__ __forceinline__ uint32_t lookup_word(const RayLane &r) { return r.lk | (r.s23 >> 16) << 20; }

// Iterations done: every iteration ends its lookup chain at exactly one level.
__device__ __forceinline__ uint32_t lane_iterations(const RayLane &r) {
    return (r.s01 & 0xFFFFu) + (r.s01 >> 16) + (r.s23 & 0xFFFFu) + (r.s23 >> 16);
}

__device__ __forceinline__ uint32_t lane_state(const RayLane &r) { return r.state; }
__device__ __forceinline__ void set_state(RayLane &r, uint32_t st) { r.state = st; }

// Node records and leaf words are addressed by 32-bit byte offsets from the pool
// base, which the hardware adds to a uniform 64-bit base in the load itself
// (wx_grid_create bounds the node pools to 4 GiB). A leaf pool beyond 4 GiB (more
// than 2^24 leaves: a solid level set of a 4096^3 domain can hold that many) is
// addressed in 64 bits instead, a grid-uniform branch.
__device__ __forceinline__ uint4 load_rec(const uint4 *pool, uint32_t byte_off) {
    uint4 q = *reinterpret_cast<const uint4 *>(reinterpret_cast<const char *>(pool) + byte_off);
    // all four words are live here, so the record stays one 16-byte load (the compiler
    // would otherwise split it into dependent partial loads inside the decode branches)
    asm volatile("" : "+v"(q.x), "+v"(q.y), "+v"(q.z), "+v"(q.w));
    return q;
}
template <int kWide = -1>
__device__ __forceinline__ uint32_t load_leaf_word(const GridDev &g, uint32_t leaf, uint32_t byte_in_leaf) {
    uint32_t v;
    if (kWide == 1 || (kWide < 0 && g.wide)) {
        v = *reinterpret_cast<const uint32_t *>(reinterpret_cast<const char *>(g.leaf) +
                                                (((uint64_t)leaf << 8) | byte_in_leaf));
        // keeps the two loads apart: merged, both would take the 64-bit address form
        asm volatile("" : "+v"(v));
    } else {
        v = *reinterpret_cast<const uint32_t *>(reinterpret_cast<const char *>(g.leaf) + ((leaf << 8) | byte_in_leaf));
    }
    return v;
}

// Workgroup -> image tile under XCD-aware dispatch. The hardware hands workgroup b
// to XCD b % 8; with `band` the tiles are renumbered so that each XCD takes one
// contiguous band of the nb row-major tiles and its own L2 holds that band's nodes
// (otherwise neighbouring tiles land on different XCDs). A bijection on [0, nb).
constexpr uint32_t kXcds = 8u;
__device__ __forceinline__ void tile_of_block(bool band, uint32_t &bx, uint32_t &by) {
    bx = blockIdx.x, by = blockIdx.y;
    if (!band) return;
    const uint32_t nb = gridDim.x * gridDim.y, b = by * gridDim.x + bx;
    const uint32_t q = nb / kXcds, r = nb % kXcds, x = b % kXcds, k = b / kXcds;
    const uint32_t t = x < r ? x * (q + 1u) + k : r * (q + 1u) + (x - r) * q + k;
    bx = t % gridDim.x, by = t / gridDim.x;
}

// Entry clip (extension of the reference's bound test, identical in the oracle):
// a ray starting outside [-bound, bound]^3 starts at its slab entry point,
// clamped into the box; a ray missing the box ends OOB with no lookups.
__device__ __forceinline__ void lane_start(RayLane &r, float ox, float oy, float oz, float dx, float dy, float dz,
                                           float B, bool clip) {
    r.dx = dx, r.dy = dy, r.dz = dz;
    r.ix = 1.0f / dx, r.iy = 1.0f / dy, r.iz = 1.0f / dz;  // :88
    // quiet a signalling NaN so that a max over |p| ignores it like the
    // comparisons of the bound test do (every other value passes unchanged)
    r.px = __builtin_canonicalizef(ox), r.py = __builtin_canonicalizef(oy), r.pz = __builtin_canonicalizef(oz);
    r.ax = 0.0f, r.ay = 0.0f, r.az = 0.0f;
    r.s01 = 0, r.s23 = 0, r.it = 0, r.lk = 0, r.np = 0, r.state = kRun;
    r.qx = 0, r.qy = 0, r.qz = 0, r.i5 = 0, r.i4 = 0, r.i3 = 0;
    if (clip && (B < fabsf(ox) || B < fabsf(oy) || B < fabsf(oz))) {
        float tn = -__builtin_inff(), tf = __builtin_inff();
        float a = (-B - ox) * r.ix, b = (B - ox) * r.ix;
        tn = smax(tn, smin(a, b));
        tf = smin(tf, smax(a, b));
        a = (-B - oy) * r.iy, b = (B - oy) * r.iy;
        tn = smax(tn, smin(a, b));
        tf = smin(tf, smax(a, b));
        a = (-B - oz) * r.iz, b = (B - oz) * r.iz;
        tn = smax(tn, smin(a, b));
        tf = smin(tf, smax(a, b));
        if (!(tn <= tf) || tf < 0.0f) {
            set_state(r, WX_OOB);
            return;
        }
        r.px = smin(smax(ox + tn * dx, -B), B);
        r.py = smin(smax(oy + tn * dy, -B), B);
        r.pz = smin(smax(oz + tn * dz, -B), B);
    }
}

// Table of RN(1/d) for d = 1..255, one per block in LDS.
__device__ __forceinline__ void init_rcp(float *s_rcp) {
    for (uint32_t i = threadIdx.x; i < 256u; i += blockDim.x) s_rcp[i] = 1.0f / (float)i;
    __syncthreads();
}

// Step scale per lookup result, one per block in LDS: entry np << 8 | d holds
// size = d * {4096, 128, 8, 1}[np] (:105-112) and RN(1/size) = RN(1/d) * 2^-shift
// (exact power-of-two scaling: no value here leaves the normal range).
constexpr uint32_t kTab = 1024u;
constexpr uint32_t kShifts = 0x0003070Cu;  // log2 of the level scales, one byte per num_parents
__device__ __forceinline__ void init_tab(float2 *s_tab) {
    for (uint32_t d = threadIdx.x; d < 256u; d += blockDim.x) {
        const float f = (float)d, rd = 1.0f / f;
#pragma unroll
        for (uint32_t np = 0; np < 4u; ++np) {
            const int32_t sh = (int32_t)((kShifts >> (np * 8u)) & 0xFFu);
            s_tab[np << 8 | d] = make_float2(ldexpf(f, sh), ldexpf(rd, -sh));
        }
    }
    __syncthreads();
}

// A ray whose direction, reciprocal direction and start are all finite never
// produces a NaN in the step: every t is a product of finite numbers (at worst an
// overflow to +-inf, never inf * 0), so the position stays free of NaN, and a
// position that overflows fails the bound test before it is stepped again. Only
// such rays take the step's NaN-free shortcuts.
__device__ __forceinline__ bool lane_clean(const RayLane &r) {
    constexpr float kInf = __builtin_inff();
    return fabsf(r.dx) < kInf && fabsf(r.dy) < kInf && fabsf(r.dz) < kInf && fabsf(r.ix) < kInf &&
           fabsf(r.iy) < kInf && fabsf(r.iz) < kInf && fabsf(r.px) < kInf && fabsf(r.py) < kInf &&
           fabsf(r.pz) < kInf;
}

// p / size, correctly rounded, for size = d * 2^k (d < 256). With y = RN(1/size)
// the Markstein sequence q0 = RN(p*y), r = p - size*q0 (exact by FMA),
// q = RN(q0 + r*y) returns the IEEE quotient whenever no intermediate
// underflows or overflows. The caller only takes this path for
// 2^-64 <= |p| <= 2^100. tests/test_fastdiv.py compares it with IEEE division
// for every divisor the layout can produce, on a stride of p's significands in
// the guard binades at both ends of that range and in two interior ones; its
// slow-marked run covers every significand of [1, 2), which covers every
// binade in range because correct rounding is scale-invariant while nothing
// under- or overflows.
__device__ __forceinline__ float div_tile(float p, float size, float y) {
    const float q0 = p * y;
    const float rr = __builtin_fmaf(-q0, size, p);
    return __builtin_fmaf(rr, y, q0);
}

// One iteration of the hdda_ray loop (raycast.comp.wgsl:92-123).
//  kN5Lds    the grid has one node-5, whose 4096 records the caller staged in LDS (s_n5);
//  kClean    every running ray of the wave is clean (lane_clean): the fused floor
//            conversion, and the axis mask as "t equals the minimum";
//  kScalarIt the caller counts the iterations and sets WX_MAXSTEPS (all lanes of the
//            wave started together, so the count is the loop's trip count); r.it is
//            not kept;
//  Tab       float: s_tab is the RN(1/d) table of init_rcp; float2: the step-scale
//            table of init_tab.
// Every combination computes the same hit record.
//  kRoot     0: the grid has one node-5, whose key is compared in scalar registers;
//            1: the root table is scanned (s_root in LDS when root_lds, else g.root);
//            -1: decided per call from g.n5count (kernels whose lanes trace different grids);
//  kWide     1 / 0: the grid's leaf pool takes 64-bit / 32-bit addresses; -1: from g.wide.
//  s_root, root_lds  the workgroup's LDS copy of the root table (stage_root) and whether
//            it holds the table.
template <bool kN5Lds = false, bool kClean = false, bool kScalarIt = false, typename Tab = float, int kRoot = -1,
          int kWide = -1>
__device__ __forceinline__ void lane_step(const GridDev &g, RayLane &r, const Tab *s_tab,
                                          const uint4 *s_n5 = nullptr, const int4 *s_root = nullptr,
                                          bool root_lds = false) {
    int32_t x, y, z;
    if (kClean) {
        x = f2i_floor(r.px), y = f2i_floor(r.py), z = f2i_floor(r.pz);
    } else {
        x = f2i(floorf(r.px)), y = f2i(floorf(r.py)), z = f2i(floorf(r.pz));
    }
    uint32_t np = r.np;
    // get_vdb_leaf_from_leaf (:360-396): deepest cached node still containing pos, i.e.
    // min(num_parents, 3 if diff < 8, 2 if < 128, 1 if < 4096, else 0). The class of
    // diff comes from its leading zeros h (v_ffbh: -1 for 0): 2-bit entries of
    // kLevelOfClz for h = 16..31 (h clamped there; h >= 29 is 3, >= 25 is 2, >= 20 is 1)
    const uint32_t diff = (uint32_t)((x ^ r.qx) | (y ^ r.qy) | (z ^ r.qz));
    uint32_t h;
    asm("v_ffbh_u32 %0, %1" : "=v"(h) : "v"(diff));
    h = h < 16u ? 16u : (h > 31u ? 31u : h);
    constexpr uint32_t kLevelOfClz = 0xFEA95500u;
    const uint32_t cls = (kLevelOfClz >> ((h << 1) - 32u)) & 3u;
    uint32_t lvl = cls < np ? cls : np;
    r.qx = x, r.qy = y, r.qz = z;
    uint32_t dist = 0;
    bool done = false;
    if (lvl == 0) {  // get_vdb_leaf_from_nothing (:398-413): the first root entry with this key
        const int32_t kx = x >> 12, ky = y >> 12, kz = z >> 12;
        int32_t found = -1;
        if (kRoot == 0 || (kRoot < 0 && g.n5count == 1u)) {  // the key in scalar registers
            if (kx == g.root0.x && ky == g.root0.y && kz == g.root0.z) found = g.root0.w;
        } else if (root_lds) {
            for (uint32_t k = 0; k < g.n5count; ++k) {
                const int4 e = s_root[k];
                if (e.x == kx && e.y == ky && e.z == kz) {
                    found = e.w;
                    break;
                }
            }
        } else {
            for (uint32_t k = 0; k < g.n5count; ++k) {
                const int4 e = g.root[k];
                if (e.x == kx && e.y == ky && e.z == kz) {
                    found = e.w;
                    break;
                }
            }
        }
        // a miss resolves at depth 0 with dist 1 (:412); a hit continues at node-5. Selects,
        // not branches: i5 is only read below num_parents 1, so a miss may overwrite it
        const bool miss = found < 0;
        dist = miss ? 1u : 0u;
        np = miss ? 0u : np;
        done = miss;
        r.s01 += miss ? 1u : 0u;
        r.i5 = (uint32_t)found;
        lvl = miss ? 0u : 1u;
    }
    if (!done && lvl == 1) {  // get_vdb_leaf_from_node5 (:415-444)
        r.lk += 1u;
        // record (off >> 3) of 4096 16-byte records per node, off = x<<10 | y<<5 | z
        const uint32_t rec = (((uint32_t)x >> 7) & 31u) << 7 | (((uint32_t)y >> 7) & 31u) << 2 |
                             (((uint32_t)z >> 7) & 31u) >> 3;
        const uint4 q = kN5Lds ? s_n5[rec] : load_rec(g.n5, (r.i5 << 16) | (rec << 4));
        const uint32_t b = ((uint32_t)z >> 7) & 7u;
        np = 1;
        // vals before kids (:424-429): an active tile's byte is 0, and a child slot that
        // is also an active tile is not walked into
        if ((q.x >> (8u + b)) & 1u) {
            r.i4 = q.y + (uint32_t)__builtin_popcount(q.x & ((1u << b) - 1u));
            lvl = 2;
        } else {
            dist = (uint32_t)((((uint64_t)q.w << 32) | q.z) >> (b * 8u)) & 0xFFu;
            done = true;
            r.s01 += 1u << 16;
        }
    }
    if (!done && lvl == 2) {  // get_vdb_leaf_from_node4 (:446-475)
        r.lk += 1u << 10;
        // record (off >> 4) of 256 per node, off = x<<8 | y<<4 | z
        const uint32_t rec = (((uint32_t)x >> 3) & 15u) << 4 | (((uint32_t)y >> 3) & 15u);
        const uint4 q = load_rec(g.n4, (r.i4 << 12) | (rec << 4));
        const uint32_t b = ((uint32_t)z >> 3) & 15u;
        np = 2;
        if ((q.x >> (16u + b)) & 1u) {
            r.i3 = q.y + (uint32_t)__builtin_popcount(q.x & ((1u << b) - 1u));
            lvl = 3;
        } else {
            dist = (uint32_t)((((uint64_t)q.w << 32) | q.z) >> (b * 4u)) & 15u;
            done = true;
            r.s23 += 1u;
        }
    }
    if (!done) {  // get_vdb_leaf_from_node3 (:477-494)
        r.s23 += 1u << 16;
        // word (off >> 3) of 64 per leaf holds 8 4-bit slots, off = x<<6 | y<<3 | z
        const uint32_t w = load_leaf_word<kWide>(g, r.i3, ((uint32_t)x & 7u) << 5 | ((uint32_t)y & 7u) << 2);
        dist = (w >> (((uint32_t)z & 7u) * 4u)) & 15u;
        np = 3;
    }
    if (!kScalarIt) r.it += 1u;
    r.np = np;
    // :96-98 hit, then :101-103 any(bound < |p|): the largest |p| exceeds the bound
    // (NaN components compare false either way); one exit for both
    const float hi = fmaxf(fmaxf(fabsf(r.px), fabsf(r.py)), fabsf(r.pz));
    if (dist == 0u || g.bound < hi) {
        r.state = dist == 0u ? WX_HIT : WX_OOB;
        return;
    }
    const float sx = r.dx < 0.0f ? -1.0f : 1.0f;  // sign11 (:70-76)
    const float sy = r.dy < 0.0f ? -1.0f : 1.0f;
    const float sz = r.dz < 0.0f ? -1.0f : 1.0f;
    // size = f32(dist) * {4096, 128, 8, 1}[num_parents] (:105-112); dist < 256, so
    // scaling by the power of two with ldexp is exact and equal to the product
    constexpr bool kStepTab = sizeof(Tab) == sizeof(float2);
    float size, yr = 0.0f;  // yr = RN(1/size), for the quotient below
    int32_t sh = 0;
    if constexpr (kStepTab) {
        const float2 e = s_tab[np << 8 | dist];
        size = e.x, yr = e.y;
    } else {
        sh = (int32_t)((kShifts >> (np * 8u)) & 0xFFu);
        size = ldexpf((float)dist, sh);
    }
    // modulo_vec3f (:78-80): p - size * floor(p / size)
    float qx, qy, qz;
    // the quotient's guard 2^-64 <= |p| <= 2^100: the upper half holds here already, as
    // hi <= bound <= 2^100 (wx_grid_create) or hi is NaN and so is lo
    const float lo = fminf(fminf(fabsf(r.px), fabsf(r.py)), fabsf(r.pz));
    if constexpr (!kStepTab) yr = ldexpf(s_tab[dist], -sh);  // exact power-of-two scaling
    qx = div_tile(r.px, size, yr);
    qy = div_tile(r.py, size, yr);
    qz = div_tile(r.pz, size, yr);
    if (!(lo >= 0x1p-64f)) {  // rare: a component below 2^-64 (or NaN) takes IEEE division
        qx = r.px / size;
        qy = r.py / size;
        qz = r.pz / size;
    }
    const float mx = r.px - size * floorf(qx);
    const float my = r.py - size * floorf(qy);
    const float mz = r.pz - size * floorf(qz);
    const float tx = r.ix * (size * smax(0.0f, sx) - mx);  // :114
    const float ty = r.iy * (size * smax(0.0f, sy) - my);
    const float tz = r.iz * (size * smax(0.0f, sz) - mz);
    const float tm = smin(smin(tx, ty), tz);  // :116
    r.px = r.px + tm * r.dx;
    r.py = r.py + tm * r.dy;
    r.pz = r.pz + tm * r.dz;
    bool bx, by, bz;  // :118-120
    if (kClean) {
        // no NaN among the t: t <= both others exactly when t equals their minimum
        bx = tx == tm, by = ty == tm, bz = tz == tm;
    } else {
        bx = (tx <= ty) & (tx <= tz);
        by = (ty <= tz) & (ty <= tx);
        bz = (tz <= tx) & (tz <= ty);
    }
    // :122, p + (4e-4 * step) * f32(mask): the product is 4e-4 * step or a zero
    // with the sign of step, selected here
    const float nx = 4e-4f * sx, ny = 4e-4f * sy, nz = 4e-4f * sz;
    r.ax = bx ? nx : 0.0f * nx;
    r.ay = by ? ny : 0.0f * ny;
    r.az = bz ? nz : 0.0f * nz;
    r.px = r.px + r.ax;
    r.py = r.py + r.ay;
    r.pz = r.pz + r.az;
    if (!kScalarIt && r.it >= kMaxSteps) r.state = WX_MAXSTEPS;
}

// The loop of lane_run for one grid-uniform choice of step variant, root lookup
// and leaf addressing: at most kMaxSteps iterations, the trip count a scalar.
template <bool kN5Lds, bool kClean, int kRoot, int kWide>
__device__ __forceinline__ void lane_loop(const GridDev &g, RayLane &r, const float2 *s_tab, const uint4 *s_n5,
                                          const int4 *s_root, bool root_lds) {
    // a wave-uniform loop (the trip count stays scalar) whose body runs on the lanes
    // still stepping; it ends when no lane of the wave steps or after kMaxSteps trips
    uint32_t it = 0;
    do {
        if (lane_state(r) == kRun)
            lane_step<kN5Lds, kClean, true, float2, kRoot, kWide>(g, r, s_tab, s_n5, s_root, root_lds);
        ++it;
    } while (it < kMaxSteps && __builtin_amdgcn_ballot_w64(lane_state(r) == kRun) != 0ull);
    if (lane_state(r) == kRun) set_state(r, WX_MAXSTEPS);
}

// Runs the rays of a wave whose lanes all started together (the one-lane-per-ray
// shapes) to their end (WX_HIT, WX_OOB, or WX_MAXSTEPS after kMaxSteps
// iterations). The iteration count is the loop's trip count, a scalar; a wave
// whose running rays are all clean takes the NaN-free step, any other wave the
// general one. The grid-uniform choices (one node-5 or a root table, 32- or 64-bit
// leaf addresses) are made once here, so the loop carries no branch on them.
template <bool kN5Lds = false>
__device__ __forceinline__ void lane_run(const GridDev &g, RayLane &r, const float2 *s_tab,
                                         const uint4 *s_n5 = nullptr, const int4 *s_root = nullptr,
                                         bool root_lds = false) {
    const bool clean = __ballot(lane_state(r) == kRun && !lane_clean(r)) == 0ull;
    const bool one = g.n5count == 1u, wide = g.wide != 0u;
    if (clean) {
        if (one) {
            if (wide) lane_loop<kN5Lds, true, 0, 1>(g, r, s_tab, s_n5, s_root, root_lds);
            else lane_loop<kN5Lds, true, 0, 0>(g, r, s_tab, s_n5, s_root, root_lds);
        } else {
            if (wide) lane_loop<kN5Lds, true, 1, 1>(g, r, s_tab, s_n5, s_root, root_lds);
            else lane_loop<kN5Lds, true, 1, 0>(g, r, s_tab, s_n5, s_root, root_lds);
        }
    } else {
        if (one) {
            if (wide) lane_loop<kN5Lds, false, 0, 1>(g, r, s_tab, s_n5, s_root, root_lds);
            else lane_loop<kN5Lds, false, 0, 0>(g, r, s_tab, s_n5, s_root, root_lds);
        } else {
            if (wide) lane_loop<kN5Lds, false, 1, 1>(g, r, s_tab, s_n5, s_root, root_lds);
            else lane_loop<kN5Lds, false, 1, 0>(g, r, s_tab, s_n5, s_root, root_lds);
        }
    }
}

// cp_main (raycast.comp.wgsl:62-63): dir = normalize(p.x*u + p.y*mv + wp), p = pixel + offset,
// with the camera words u, mv, wp (ComputeState bytes 144..191).
__device__ __forceinline__ void camera_ray(const float *u, const float *mv, const float *wp, float off, uint32_t px,
                                           uint32_t py, float &dx, float &dy, float &dz) {
    const float fx = (float)px + off, fy = (float)py + off;
    const float vx = (fx * u[0] + fy * mv[0]) + wp[0];
    const float vy = (fx * u[1] + fy * mv[1]) + wp[1];
    const float vz = (fx * u[2] + fy * mv[2]) + wp[2];
    const float len = sqrtf((vx * vx + vy * vy) + vz * vz);
    dx = vx / len, dy = vy / len, dz = vz / len;
}

}  // namespace wxk

// A device-resident grid (the opaque wx_grid of the C ABI): the four pools and
// their device addresses.
struct wx_grid {
    wxk::GridDev dev;
    void *mem[4] = {nullptr, nullptr, nullptr, nullptr};
    uint64_t bytes = 0;
    int device = 0;
    uint32_t n5 = 0, n4 = 0, n3 = 0;  // node counts of the layout
};
