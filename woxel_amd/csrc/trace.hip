// HDDA + grid-SDF traversal kernels for gfx950, and the device-side C ABI.
//
// The per-lane traversal (RayLane, lane_step) lives in lane.hpp. The launch shapes
// (wx_tune.persistent) share it:
//  * k_simple_*: one lane per ray, 8x8-pixel wave tiles (the reference's dispatch
//    shape); camera tiles dealt round-robin over the XCDs, or in one band per XCD
//    (wx_tune.xcd_split = 2: better L2 locality, worse balance -- measured slower);
//  * k_persist*: persistent wave64 workers. A wave holds 64 rays; once `refill`
//    lanes have finished, the idle lanes take the next slots of the wave's chunk
//    of the work queue (chunks come from per-XCD queue regions, with stealing).
//    Shadow rays continue in the lane of their primary. Every wave exits once
//    the queue is drained and its lanes are idle; every ray ends within 1000
//    steps, so the grid always drains;
//  * k_compact_camera: 16x16-pixel workgroups that move the rays of their emptiest
//    wave into idle lanes of the others every `refill` iterations;
//  * k_lds_camera: one lane per ray with the grid's node-5 records staged in LDS.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "lane.hpp"
#include "request.hpp"
#include "wx_internal.hpp"

using wx::set_error;

namespace {

using namespace wxk;

constexpr uint32_t kRegions = 8u;      // queue regions (one per XCD)
constexpr uint32_t kMaxGridsLds = 64u; // grid table entries staged in LDS
constexpr uint32_t kMaxReqsLds = 1024u;

// ---------------------------------------------------------------- simple shape

// Shadow continuation of a finished primary in the same lane (WX_TRACE_SHADOW).
template <bool kN5Lds = false>
__device__ __forceinline__ void finish_with_shadow(const GridDev &g, const ReqDev &q, RayLane &r, uint32_t pix,
                                                   const float2 *s_tab, const uint4 *s_n5 = nullptr,
                                                   const int4 *s_root = nullptr, bool root_lds = false) {
    if (lane_state(r) != WX_HIT) {
        write_no_ray(q.shadow + pix);
        return;
    }
    float ox, oy, oz;
    shadow_origin(r, ox, oy, oz);
    start_shadow(r, q, g.bound, ox, oy, oz);
    lane_run<kN5Lds>(g, r, s_tab, s_n5, s_root, root_lds);
    lane_write(r, ox, oy, oz, q.shadow + pix);
}

// 128-lane workgroups of two 8x8 wave tiles (16x8 pixels): a finished workgroup frees
// its slots for the next one sooner than a 16x16 one (measured faster).
template <bool kShadow>
__global__ __launch_bounds__(128) void k_simple_camera(GridDev g, ReqDev q, uint32_t band) {
    __shared__ float2 s_tab[kTab];
    __shared__ int4 s_root_buf[kRootLds];
    const bool root_lds = stage_root(g, s_root_buf);
    init_tab(s_tab);
    uint32_t bx, by;
    tile_of_block(band != 0u, bx, by);
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t px = bx * 16u + wave * 8u + (lane & 7u);
    const uint32_t py = by * 8u + (lane >> 3);
    if (px >= q.w || py >= q.h) return;
    float dx, dy, dz;
    camera_dir(q, px, py, dx, dy, dz);
    RayLane r;
    lane_start(r, q.ex, q.ey, q.ez, dx, dy, dz, g.bound, (q.flags & WX_TRACE_CLIP) != 0u);
    lane_run(g, r, s_tab, nullptr, s_root_buf, root_lds);
    const uint32_t pix = py * q.w + px;
    lane_write(r, q.ex, q.ey, q.ez, q.out + pix);
    if (kShadow) finish_with_shadow(g, q, r, pix, s_tab, nullptr, s_root_buf, root_lds);
}

// The one-lane shape with the grid's single node-5 (4096 records, 64 KiB) staged in
// LDS: a 1024-lane workgroup covers 32x32 pixels (sixteen 8x8 wave tiles), so two
// workgroups (32 waves) share a CU's LDS. Node-5 lookups read LDS instead of the
// cache hierarchy; everything else is k_simple_camera.
template <bool kShadow>
__global__ __launch_bounds__(1024) void k_lds_camera(GridDev g, ReqDev q) {
    __shared__ float2 s_tab[kTab];
    __shared__ uint4 s_n5[4096];
    for (uint32_t i = threadIdx.x; i < 4096u; i += 1024u) s_n5[i] = g.n5[i];
    init_tab(s_tab);  // its barrier also publishes s_n5
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t px = blockIdx.x * 32u + (wave & 3u) * 8u + (lane & 7u);
    const uint32_t py = blockIdx.y * 32u + (wave >> 2) * 8u + (lane >> 3);
    if (px >= q.w || py >= q.h) return;
    float dx, dy, dz;
    camera_dir(q, px, py, dx, dy, dz);
    RayLane r;
    lane_start(r, q.ex, q.ey, q.ez, dx, dy, dz, g.bound, (q.flags & WX_TRACE_CLIP) != 0u);
    lane_run<true>(g, r, s_tab, s_n5);
    const uint32_t pix = py * q.w + px;
    lane_write(r, q.ex, q.ey, q.ez, q.out + pix);
    if (kShadow) finish_with_shadow<true>(g, q, r, pix, s_tab, s_n5);
}

template <bool kShadow>
__global__ __launch_bounds__(128) void k_simple_rays(GridDev g, ReqDev q) {
    __shared__ float2 s_tab[kTab];
    __shared__ int4 s_root_buf[kRootLds];
    const bool root_lds = stage_root(g, s_root_buf);
    init_tab(s_tab);
    const uint32_t i = blockIdx.x * 128u + threadIdx.x;
    if (i >= q.w) return;
    const float *p = reinterpret_cast<const float *>(q.rays + i);
    RayLane r;
    lane_start(r, p[0], p[1], p[2], p[3], p[4], p[5], g.bound, (q.flags & WX_TRACE_CLIP) != 0u);
    lane_run(g, r, s_tab, nullptr, s_root_buf, root_lds);
    lane_write(r, p[0], p[1], p[2], q.out + i);
    if (kShadow) finish_with_shadow(g, q, r, i, s_tab, nullptr, s_root_buf, root_lds);
}

// ---------------------------------------------------------------- compact shape

// A camera request (primary rays only) traced by 256-lane workgroups of 16x16
// pixels that keep their lanes busy: every `K` iterations the workgroup counts
// the rays still running in each of its four waves; when the emptiest wave's
// rays fit into the idle lanes of the other waves that still run, they move
// there through LDS (the whole lane state, so results do not change) and that
// wave idles. Every lane stays in the loop until the workgroup has no ray left
// (the barriers need all 256), and every round advances each running ray by at
// least one iteration, so the loop ends.
constexpr int kLaneWords = 25;

__device__ __forceinline__ void put_lane(uint32_t *buf, uint32_t slot, const RayLane &r, uint32_t pix) {
    const uint32_t w[kLaneWords] = {
        __float_as_uint(r.dx), __float_as_uint(r.dy), __float_as_uint(r.dz), __float_as_uint(r.ix),
        __float_as_uint(r.iy), __float_as_uint(r.iz), __float_as_uint(r.px), __float_as_uint(r.py),
        __float_as_uint(r.pz), __float_as_uint(r.ax), __float_as_uint(r.ay), __float_as_uint(r.az),
        r.s01, r.s23, r.it, r.lk, r.np, r.state, (uint32_t)r.qx, (uint32_t)r.qy,
        (uint32_t)r.qz, r.i5, r.i4, r.i3, pix};
#pragma unroll
    for (int k = 0; k < kLaneWords; ++k) buf[k * 64 + slot] = w[k];
}

__device__ __forceinline__ void get_lane(const uint32_t *buf, uint32_t slot, RayLane &r, uint32_t &pix) {
    uint32_t w[kLaneWords];
#pragma unroll
    for (int k = 0; k < kLaneWords; ++k) w[k] = buf[k * 64 + slot];
    r.dx = __uint_as_float(w[0]), r.dy = __uint_as_float(w[1]), r.dz = __uint_as_float(w[2]);
    r.ix = __uint_as_float(w[3]), r.iy = __uint_as_float(w[4]), r.iz = __uint_as_float(w[5]);
    r.px = __uint_as_float(w[6]), r.py = __uint_as_float(w[7]), r.pz = __uint_as_float(w[8]);
    r.ax = __uint_as_float(w[9]), r.ay = __uint_as_float(w[10]), r.az = __uint_as_float(w[11]);
    r.s01 = w[12], r.s23 = w[13];
    r.it = w[14], r.lk = w[15], r.np = w[16], r.state = w[17];
    r.qx = (int32_t)w[18], r.qy = (int32_t)w[19], r.qz = (int32_t)w[20];
    r.i5 = w[21], r.i4 = w[22], r.i3 = w[23];
    pix = w[24];
}

__global__ __launch_bounds__(256) void k_compact_camera(GridDev g, ReqDev q, uint32_t K) {
    __shared__ float s_rcp[256];
    __shared__ uint32_t s_buf[kLaneWords * 64];
    __shared__ uint32_t s_cnt[4];
    __shared__ int4 s_root_buf[kRootLds];
    const bool root_lds = stage_root(g, s_root_buf);
    init_rcp(s_rcp);
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t px = blockIdx.x * 16u + (wave & 1u) * 8u + (lane & 7u);
    const uint32_t py = blockIdx.y * 16u + (wave >> 1) * 8u + (lane >> 3);
    const uint64_t below = (1ull << lane) - 1ull;
    RayLane r;
    r.state = 0;
    uint32_t pix = 0;
    bool alive = false;
    if (px < q.w && py < q.h) {
        float dx, dy, dz;
        camera_dir(q, px, py, dx, dy, dz);
        lane_start(r, q.ex, q.ey, q.ez, dx, dy, dz, g.bound, (q.flags & WX_TRACE_CLIP) != 0u);
        pix = py * q.w + px;
        alive = true;
    }
    for (;;) {
        if (alive) {
            for (uint32_t k = 0; k < K && lane_state(r) == kRun; ++k)
                lane_step(g, r, s_rcp, nullptr, s_root_buf, root_lds);
            if (lane_state(r) != kRun) {
                lane_write(r, q.ex, q.ey, q.ez, q.out + pix);
                alive = false;
            }
        }
        const uint64_t am = __ballot(alive);
        if (lane == 0) s_cnt[wave] = (uint32_t)__popcll(am);
        __syncthreads();
        const uint32_t c0 = s_cnt[0], c1 = s_cnt[1], c2 = s_cnt[2], c3 = s_cnt[3];
        const uint32_t c[4] = {c0, c1, c2, c3};
        if (c0 + c1 + c2 + c3 == 0u) break;
        // the emptiest wave that still runs, if the other running waves can take its rays
        uint32_t src = 4u, running = 0u, room = 0u;
        for (uint32_t w = 0; w < 4u; ++w)
            if (c[w]) {
                ++running;
                if (src == 4u || c[w] < c[src]) src = w;
            }
        for (uint32_t w = 0; w < 4u; ++w)
            if (c[w] && w != src) room += 64u - c[w];
        if (running >= 2u && c[src] <= room) {
            if (wave == src && alive) {
                put_lane(s_buf, (uint32_t)__popcll(am & below), r, pix);
                alive = false;
            }
            __syncthreads();
            if (wave != src && c[wave] && !alive) {
                uint32_t rank = 0;  // this idle lane's place among the idle lanes of the receiving waves
                for (uint32_t w = 0; w < wave; ++w)
                    if (c[w] && w != src) rank += 64u - c[w];
                rank += (uint32_t)__popcll(~am & below);
                if (rank < c[src]) {
                    get_lane(s_buf, rank, r, pix);
                    alive = true;
                }
            }
        }
        __syncthreads();
    }
}

// Any batch, one lane per slot: slot s of the launch is request r's slot
// s - begin[r]; a primary hit with WX_TRACE_SHADOW continues as the shadow ray.
struct BatchArgs {
    const GridDev *grids;
    const ReqDev *reqs;
    const uint32_t *begin;
    uint32_t n_grids, n_reqs, total, pad;
};

__global__ __launch_bounds__(128) void k_simple_batch(BatchArgs a) {
    __shared__ float2 s_tab[kTab];
    init_tab(s_tab);
    const uint32_t slot = blockIdx.x * 128u + threadIdx.x;
    if (slot >= a.total) return;  // whole waves: every request has a multiple of 64 slots
    // so the 64 slots of a wave belong to one request: find it once per wave
    const uint32_t wslot = (uint32_t)__builtin_amdgcn_readfirstlane((int)(slot & ~63u));
    uint32_t lo = 0, hi = a.n_reqs;  // slot -> request: binary search over the prefix
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a.begin[mid] <= wslot) lo = mid;
        else hi = mid;
    }
    lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)lo);
    const ReqDev &q = a.reqs[lo];
    const GridDev g = a.grids[q.grid];
    const uint32_t s = slot - a.begin[lo];
    float ox, oy, oz, dx, dy, dz;
    uint32_t pix;
    if (q.kind == WX_REQ_CAMERA) {
        uint32_t px, py;
        if (!camera_pixel(q, s, px, py)) return;
        pix = py * q.w + px;
        ox = q.ex, oy = q.ey, oz = q.ez;
        camera_dir(q, px, py, dx, dy, dz);
    } else {
        if (s >= q.w) return;
        pix = s;
        const float *p = reinterpret_cast<const float *>(q.rays + s);
        ox = p[0], oy = p[1], oz = p[2], dx = p[3], dy = p[4], dz = p[5];
    }
    RayLane r;
    lane_start(r, ox, oy, oz, dx, dy, dz, g.bound, (q.flags & WX_TRACE_CLIP) != 0u);
    lane_run(g, r, s_tab);
    lane_write(r, ox, oy, oz, q.out + pix);
    if (q.flags & WX_TRACE_SHADOW) finish_with_shadow(g, q, r, pix, s_tab);
}

// ---------------------------------------------------------------- persistent shape

struct PersistArgs {
    const GridDev *grids;  // n_grids
    const ReqDev *reqs;    // n_reqs
    const uint32_t *begin; // n_reqs + 1 slot prefix (camera: 64 slots per 8x8 tile; rays: one per ray)
    uint32_t *ctr;         // kRegions counters, 16 words apart
    uint32_t n_grids, n_reqs, total, regions;
    uint32_t refill, chunk;
    uint32_t pad[2];
};

__global__ __launch_bounds__(256) void k_zero(uint32_t *p, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += 256u) p[i] = 0u;
}

// Work queue of one persistent wave: chunks of `chunk` slots from the wave's
// queue region (regions are per XCD; a drained region passes the wave on to the
// next one). All values are wave-uniform.
struct WaveQueue {
    uint32_t region, probed, cur, end, regsize;

    __device__ __forceinline__ void init(const PersistArgs &a) {
        regsize = ((a.total + a.regions * 64u - 1u) / (a.regions * 64u)) * 64u;
        region = blockIdx.x % a.regions;  // workgroups are dispatched round-robin over the XCDs
        probed = 0;
        cur = end = 0;
    }
    __device__ __forceinline__ bool drained(const PersistArgs &a) const { return cur == end && probed >= a.regions; }
    // Take the next chunk once the current one is used up.
    __device__ __forceinline__ void fetch(const PersistArgs &a, uint32_t lane) {
        if (cur != end || probed >= a.regions) return;
        for (;;) {
            const uint32_t rs = region * regsize;
            uint32_t got = 0;
            if (lane == 0) got = atomicAdd(&a.ctr[region * 16u], a.chunk);
            got = __builtin_amdgcn_readfirstlane(got);
            const uint32_t re = std::min(rs + regsize, a.total);
            if (got < regsize && rs + got < re) {
                cur = rs + got;
                end = std::min(cur + a.chunk, re);
                return;
            }
            if (++probed >= a.regions) return;
            region = (region + 1u) % a.regions;
        }
    }
};

// One request (camera or ray list), no shadow rays: the single-request path.
// Outer loop: refill idle lanes from the wave's chunk. Inner loop: step the
// live lanes (the same body as the one-ray-per-lane kernel) until `refill`
// lanes have finished or none is left.
template <uint32_t KIND>
__global__ __launch_bounds__(256) void k_persist1(PersistArgs a, GridDev g, ReqDev q) {
    __shared__ float2 s_tab[kTab];
    __shared__ int4 s_root_buf[kRootLds];
    const bool root_lds = stage_root(g, s_root_buf);
    init_tab(s_tab);
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t lt = (1ull << lane) - 1ull;
    const bool clip = (q.flags & WX_TRACE_CLIP) != 0u;
    WaveQueue wq;
    wq.init(a);
    bool alive = false;
    uint32_t pix = 0;
    RayLane r;
    r.state = 0;
    for (;;) {
        const uint64_t idle = __ballot(!alive);
        const uint32_t nidle = (uint32_t)__popcll(idle);
        wq.fetch(a, lane);
        const uint32_t avail = wq.end - wq.cur;
        const uint32_t rank = (uint32_t)__popcll(idle & lt);
        if (!alive && rank < avail) {
            const uint32_t s = wq.cur + rank;
            if (KIND == WX_REQ_CAMERA) {
                uint32_t px, py;
                if (camera_pixel(q, s, px, py)) {
                    float dx, dy, dz;
                    camera_dir(q, px, py, dx, dy, dz);
                    lane_start(r, q.ex, q.ey, q.ez, dx, dy, dz, g.bound, clip);
                    pix = py * q.w + px;
                    alive = true;
                }
            } else if (s < q.w) {
                const float *p = reinterpret_cast<const float *>(q.rays + s);
                lane_start(r, p[0], p[1], p[2], p[3], p[4], p[5], g.bound, clip);
                pix = s;
                alive = true;
            }
        }
        wq.cur += std::min(avail, nidle);
        if (!__any(alive)) {
            if (wq.drained(a)) break;
            continue;
        }
        do {
            if (alive) {
                if (lane_state(r) == kRun) lane_step(g, r, s_tab, nullptr, s_root_buf, root_lds);
                if (lane_state(r) != kRun) {
                    if (KIND == WX_REQ_CAMERA) {
                        lane_write(r, q.ex, q.ey, q.ez, q.out + pix);
                    } else {
                        const float *p = reinterpret_cast<const float *>(q.rays + pix);
                        lane_write(r, p[0], p[1], p[2], q.out + pix);
                    }
                    alive = false;
                }
            }
        } while ((uint32_t)__popcll(__ballot(!alive)) < a.refill);
    }
}

// Any batch: many requests and grids, shadow rays continued in the lane of
// their primary. Ray origins live in LDS (one per lane). A lane keeps its
// request's grid index in a register and reads that grid's descriptor from LDS
// on each step; the 128-byte request record is read when a ray starts or ends.
__global__ __launch_bounds__(256) void k_persist_batch(PersistArgs a) {
    __shared__ float2 s_tab[kTab];
    __shared__ GridDev s_grid[kMaxGridsLds];
    __shared__ uint32_t s_begin[kMaxReqsLds + 1];
    __shared__ float s_org[3][256];
    for (uint32_t i = threadIdx.x; i < a.n_grids; i += 256u) s_grid[i] = a.grids[i];
    for (uint32_t i = threadIdx.x; i <= a.n_reqs; i += 256u) s_begin[i] = a.begin[i];
    init_tab(s_tab);
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t lt = (1ull << lane) - 1ull;
    WaveQueue wq;
    wq.init(a);
    bool alive = false, shadow = false;
    uint32_t req = 0, gi = 0, pix = 0;
    RayLane r;
    r.state = 0;
    for (;;) {
        const uint64_t idle = __ballot(!alive);
        const uint32_t nidle = (uint32_t)__popcll(idle);
        if (nidle >= a.refill) {
            wq.fetch(a, lane);
            const uint32_t avail = wq.end - wq.cur;
            const uint32_t rank = (uint32_t)__popcll(idle & lt);
            if (!alive && rank < avail) {
                const uint32_t slot = wq.cur + rank;
                uint32_t lo = 0, hi = a.n_reqs;  // slot -> request: binary search over the prefix
                while (hi - lo > 1u) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (s_begin[mid] <= slot) lo = mid;
                    else hi = mid;
                }
                const ReqDev &q = a.reqs[lo];
                const uint32_t s = slot - s_begin[lo];
                bool ok;
                float ox, oy, oz, dx, dy, dz;
                if (q.kind == WX_REQ_CAMERA) {
                    uint32_t px, py;
                    ok = camera_pixel(q, s, px, py);
                    pix = py * q.w + px;
                    ox = q.ex, oy = q.ey, oz = q.ez;
                    if (ok) camera_dir(q, px, py, dx, dy, dz);
                } else {
                    ok = s < q.w;
                    pix = s;
                    if (ok) {
                        const float *p = reinterpret_cast<const float *>(q.rays + s);
                        ox = p[0], oy = p[1], oz = p[2], dx = p[3], dy = p[4], dz = p[5];
                    }
                }
                if (ok) {
                    gi = q.grid;
                    lane_start(r, ox, oy, oz, dx, dy, dz, s_grid[gi].bound, (q.flags & WX_TRACE_CLIP) != 0u);
                    s_org[0][threadIdx.x] = ox, s_org[1][threadIdx.x] = oy, s_org[2][threadIdx.x] = oz;
                    req = lo;
                    shadow = false;
                    alive = true;
                }
            }
            wq.cur += std::min(avail, nidle);
        }
        if (!__any(alive)) {
            if (wq.drained(a)) break;
            continue;
        }
        if (alive) {
            const GridDev &g = s_grid[gi];
            if (lane_state(r) == kRun) lane_step(g, r, s_tab);
            if (lane_state(r) != kRun) {
                const ReqDev &q = a.reqs[req];
                const float ox = s_org[0][threadIdx.x], oy = s_org[1][threadIdx.x], oz = s_org[2][threadIdx.x];
                if (shadow) {
                    lane_write(r, ox, oy, oz, q.shadow + pix);
                    alive = false;
                } else {
                    lane_write(r, ox, oy, oz, q.out + pix);
                    alive = false;
                    if (q.flags & WX_TRACE_SHADOW) {
                        if (lane_state(r) == WX_HIT) {
                            float sx, sy, sz;
                            shadow_origin(r, sx, sy, sz);
                            start_shadow(r, q, g.bound, sx, sy, sz);
                            s_org[0][threadIdx.x] = sx, s_org[1][threadIdx.x] = sy, s_org[2][threadIdx.x] = sz;
                            shadow = true;
                            alive = true;
                            if (lane_state(r) != kRun) {  // the shadow ray missed the bound box
                                lane_write(r, sx, sy, sz, q.shadow + pix);
                                alive = false;
                            }
                        } else {
                            write_no_ray(q.shadow + pix);
                        }
                    }
                }
            }
        }
    }
}

int hip_fail(hipError_t e, const char *what) {
    return set_error(WX_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// ---------------------------------------------------------------- per-stream scratch

// Device scratch for one stream: queue counters + the batch tables, and a pinned
// host staging buffer for the tables. Launches on one stream are ordered, so a
// stream's scratch is never used by two launches at once (callers that share a
// stream between threads serialise their calls on it, as for any stream).
struct StreamScratch {
    void *dev = nullptr;
    size_t dev_cap = 0;
    void *host = nullptr;
    size_t host_cap = 0;
    hipEvent_t copied = nullptr;  // the last table upload from `host` is done
    bool pending = false;
};

std::mutex g_scratch_mu;
std::map<std::pair<int, void *>, StreamScratch> g_scratch;

int scratch_for(hipStream_t s, size_t dev_bytes, size_t host_bytes, StreamScratch **out) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    StreamScratch &S = g_scratch[{dev, (void *)s}];
    if (S.pending) {
        hipError_t e = hipEventSynchronize(S.copied);
        if (e != hipSuccess) return hip_fail(e, "scratch: event sync");
        S.pending = false;
    }
    if (S.dev_cap < dev_bytes) {
        if (S.dev) {
            hipError_t e = hipStreamSynchronize(s);
            if (e != hipSuccess) return hip_fail(e, "scratch: stream sync");
            (void)hipFree(S.dev);
        }
        size_t cap = std::max<size_t>(dev_bytes, 64u << 10);
        hipError_t e = hipMalloc(&S.dev, cap);
        if (e != hipSuccess) {
            S.dev = nullptr;
            S.dev_cap = 0;
            return set_error(WX_ERR_BUDGET, std::string("scratch: hipMalloc: ") + hipGetErrorString(e));
        }
        S.dev_cap = cap;
    }
    if (S.host_cap < host_bytes) {
        if (S.host) (void)hipHostFree(S.host);
        size_t cap = std::max<size_t>(host_bytes, 64u << 10);
        hipError_t e = hipHostMalloc(&S.host, cap, hipHostMallocDefault);
        if (e != hipSuccess) {
            S.host = nullptr;
            S.host_cap = 0;
            return set_error(WX_ERR_BUDGET, std::string("scratch: hipHostMalloc: ") + hipGetErrorString(e));
        }
        S.host_cap = cap;
    }
    if (!S.copied) {
        hipError_t e = hipEventCreateWithFlags(&S.copied, hipEventDisableTiming);
        if (e != hipSuccess) return hip_fail(e, "scratch: event");
    }
    *out = &S;
    return WX_OK;
}

}  // namespace

// Frees the scratch of a stream about to be destroyed (its work is waited for).
void wx::drop_stream_scratch(void *stream) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    for (auto it = g_scratch.begin(); it != g_scratch.end();) {
        if (it->first.second != stream) {
            ++it;
            continue;
        }
        StreamScratch &S = it->second;
        (void)hipStreamSynchronize(static_cast<hipStream_t>(stream));
        if (S.dev) (void)hipFree(S.dev);
        if (S.host) (void)hipHostFree(S.host);
        if (S.copied) (void)hipEventDestroy(S.copied);
        it = g_scratch.erase(it);
    }
}

namespace {

// ---------------------------------------------------------------- tuning

std::mutex g_tune_mu;
std::map<int, wx_tune> g_tune;

wx_tune default_tune() {
    wx_tune t;
    t.persistent = 0;
    t.refill = 16;
    t.chunk = 256;
    t.waves_per_cu = 16;
    t.xcd_split = 1;
    return t;
}

wx_tune current_tune() {
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(dev);
    return it == g_tune.end() ? default_tune() : it->second;
}

bool tune_ok(const wx_tune &t) {
    return t.persistent <= 3 && t.refill >= 1 && t.refill <= 64 && t.chunk >= 64 && t.chunk % 64 == 0 && t.chunk <= (1u << 20) &&
           t.waves_per_cu >= 1 && t.waves_per_cu <= 32 && t.xcd_split <= 2;
}

// CUs of the current device (queried once per device; thread-safe).
int cu_count() {
    static std::mutex mu;
    static std::map<int, int> cus;
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lk(mu);
    auto it = cus.find(dev);
    if (it != cus.end()) return it->second;
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
    cus[dev] = n;
    return n;
}

}  // namespace


namespace {

// Host request -> device request; checks everything the kernel indexes with.
int make_req(const wx_request &in, uint32_t n_grids, ReqDev &q, uint32_t &slots) {
    std::memset(&q, 0, sizeof q);
    if (in.grid >= n_grids) return set_error(WX_ERR_INVALID, "request: grid index out of range");
    if (in.kind != WX_REQ_CAMERA && in.kind != WX_REQ_RAYS) return set_error(WX_ERR_INVALID, "request: bad kind");
    if (in.flags & ~(WX_TRACE_CLIP | WX_TRACE_SHADOW)) return set_error(WX_ERR_INVALID, "request: unknown flags");
    if (!in.out && in.width && in.height) return set_error(WX_ERR_INVALID, "request: NULL output buffer");
    if ((in.flags & WX_TRACE_SHADOW) && !in.shadow_out && in.width && in.height)
        return set_error(WX_ERR_INVALID, "request: SHADOW without a shadow output buffer");
    q.kind = in.kind;
    q.grid = in.grid;
    q.flags = in.flags;
    q.out = in.out;
    q.shadow = in.shadow_out;
    q.lx = in.light[0], q.ly = in.light[1], q.lz = in.light[2];
    if (in.kind == WX_REQ_CAMERA) {
        if (in.width > 65536u || in.height > 65536u) return set_error(WX_ERR_INVALID, "request: image too large");
        q.w = in.width;
        q.h = in.height;
        q.tiles_x = (in.width + 7u) / 8u;
        const uint64_t tiles = (uint64_t)q.tiles_x * ((in.height + 7u) / 8u);
        if (tiles * 64u > 0x7FFFFFFFull) return set_error(WX_ERR_INVALID, "request: image too large");
        slots = (uint32_t)(tiles * 64u);
        q.off = in.pixel_offset;
        q.ex = in.camera.eye[0], q.ey = in.camera.eye[1], q.ez = in.camera.eye[2];
        q.ux = in.camera.u[0], q.uy = in.camera.u[1], q.uz = in.camera.u[2];
        q.mx = in.camera.mv[0], q.my = in.camera.mv[1], q.mz = in.camera.mv[2];
        q.wx = in.camera.wp[0], q.wy = in.camera.wp[1], q.wz = in.camera.wp[2];
        if (!in.width || !in.height) slots = 0;
    } else {
        if (!in.rays && in.width) return set_error(WX_ERR_INVALID, "request: NULL ray buffer");
        if (in.width > 0x7FFFFFC0u) return set_error(WX_ERR_INVALID, "request: too many rays");
        q.w = in.width;
        q.h = 1;
        q.rays = in.rays;
        slots = (in.width + 63u) & ~63u;
    }
    return WX_OK;
}

// Table layout of a batched launch in its scratch memory: queue counters, then the
// grid table, the request table and the slot prefix, packed.
constexpr size_t kOffCtr = 0, kOffGrids = 1024;
constexpr size_t off_reqs(uint32_t n_grids) { return kOffGrids + sizeof(GridDev) * n_grids; }
constexpr size_t off_begin(uint32_t n_grids, uint32_t n_reqs) { return off_reqs(n_grids) + sizeof(ReqDev) * n_reqs; }
constexpr size_t table_bytes(uint32_t n_grids, uint32_t n_reqs) {
    return off_begin(n_grids, n_reqs) + 4u * (n_reqs + 1u);
}

// Allocation-free launches (wx_trace_batch_static) keep the same tables in a
// file-scope device array, one copy per GPU, sized for the largest batch a launch
// takes. Host code reaches it through the address hipGetSymbolAddress returns on
// each device, never through the array's name.
constexpr size_t kStaticBytes = (table_bytes(kMaxGridsLds, kMaxReqsLds) + 255u) & ~size_t(255);
__device__ __align__(256) unsigned char scratch[kStaticBytes];

constexpr int kMaxStaticDevices = 16;
std::mutex g_static_mu;
void *g_static_dev[kMaxStaticDevices];  // scratch's address on each device (nullptr until first use)
// Host staging of the tables, one area per device. The tables are copied with
// hipMemcpyAsync from this pageable memory, which the runtime has consumed when
// the call returns, so the next call on the device may overwrite it.
alignas(256) unsigned char g_static_host[kMaxStaticDevices][kStaticBytes];

// The current device's static scratch (device address, host staging).
int static_scratch(char **dev_out, char **host_out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return set_error(WX_ERR_DEVICE, std::string("static scratch: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxStaticDevices)
        return set_error(WX_ERR_UNSUPPORTED, "static scratch: device ordinal beyond 15");
    std::lock_guard<std::mutex> lk(g_static_mu);
    if (!g_static_dev[dev]) {
        void *ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return set_error(WX_ERR_DEVICE, std::string("static scratch: hipGetSymbolAddress: ") + hipGetErrorString(e));
        g_static_dev[dev] = ptr;
    }
    *dev_out = static_cast<char *>(g_static_dev[dev]);
    *host_out = reinterpret_cast<char *>(g_static_host[dev]);
    return WX_OK;
}

// Where a launch's tables and counters live: per-stream scratch that grows on
// demand (wx_trace_batch) or the device's static scratch (wx_trace_batch_static).
struct Staging {
    bool is_static = false;
    char *host = nullptr;  // static: host staging area; tables are built in place
    char *dev = nullptr;   // static: device scratch
};

// Copies the grid table and the request/prefix tables (already at their packed
// offsets in `host`) to the device scratch `dev` and returns their device addresses.
int copy_tables(wx_grid *const *grids, uint32_t n_grids, uint32_t n_reqs, char *host, char *dev, hipStream_t stream,
                const GridDev **g_out, const ReqDev **q_out, const uint32_t **b_out) {
    for (uint32_t i = 0; i < n_grids; ++i)
        std::memcpy(host + kOffGrids + sizeof(GridDev) * i, &grids[i]->dev, sizeof(GridDev));
    const size_t bytes = table_bytes(n_grids, n_reqs);
    hipError_t e = hipMemcpyAsync(dev + kOffGrids, host + kOffGrids, bytes - kOffGrids, hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) return hip_fail(e, "batch: table upload");
    *g_out = reinterpret_cast<const GridDev *>(dev + kOffGrids);
    *q_out = reinterpret_cast<const ReqDev *>(dev + off_reqs(n_grids));
    *b_out = reinterpret_cast<const uint32_t *>(dev + off_begin(n_grids, n_reqs));
    return WX_OK;
}

// Upload the grid/request/prefix tables of a batch into the stream's scratch
// (after the counter block) and return their device addresses.
int upload_tables(wx_grid *const *grids, uint32_t n_grids, const ReqDev *q, const uint32_t *begin, uint32_t n_reqs,
                  hipStream_t stream, StreamScratch **S_out, const GridDev **g_out, const ReqDev **q_out,
                  const uint32_t **b_out) {
    const size_t bytes = table_bytes(n_grids, n_reqs);
    StreamScratch *S = nullptr;
    int rc = scratch_for(stream, bytes, bytes, &S);
    if (rc) return rc;
    char *h = static_cast<char *>(S->host);
    std::memcpy(h + off_reqs(n_grids), q, sizeof(ReqDev) * n_reqs);
    std::memcpy(h + off_begin(n_grids, n_reqs), begin, 4u * (n_reqs + 1u));
    rc = copy_tables(grids, n_grids, n_reqs, h, static_cast<char *>(S->dev), stream, g_out, q_out, b_out);
    if (rc) return rc;
    if (hipEventRecord(S->copied, stream) != hipSuccess) return set_error(WX_ERR_DEVICE, "batch: event record");
    S->pending = true;
    *S_out = S;
    return WX_OK;
}

// One launch over a validated batch, in the shape `t` selects. With a static
// staging the tables are built directly in the device's host staging area and no
// memory is allocated.
int launch_batch(wx_grid *const *grids, uint32_t n_grids, const wx_request *reqs, uint32_t n_reqs,
                 const wx_tune &t, hipStream_t stream, const Staging &stg = Staging()) {
    if (!tune_ok(t)) return set_error(WX_ERR_INVALID, "tune: out of range");
    std::vector<ReqDev> qv;
    std::vector<uint32_t> bv;
    ReqDev *q;
    uint32_t *begin;
    if (stg.is_static) {
        if (n_grids > kMaxGridsLds || n_reqs > kMaxReqsLds)
            return set_error(WX_ERR_INVALID, "batch: at most 64 grids and 1024 requests per allocation-free call");
        q = reinterpret_cast<ReqDev *>(stg.host + off_reqs(n_grids));
        begin = reinterpret_cast<uint32_t *>(stg.host + off_begin(n_grids, n_reqs));
    } else {
        qv.resize(n_reqs);
        bv.resize(n_reqs + 1u);
        q = qv.data();
        begin = bv.data();
    }
    uint64_t total = 0;
    for (uint32_t i = 0; i < n_reqs; ++i) {
        uint32_t slots = 0;
        int rc = make_req(reqs[i], n_grids, q[i], slots);
        if (rc) return rc;
        begin[i] = (uint32_t)total;
        total += slots;
        if (total > 0x7FFFFFFFull) return set_error(WX_ERR_INVALID, "batch: more than 2^31 ray slots");
    }
    begin[n_reqs] = (uint32_t)total;
    if (total == 0) return WX_OK;
    for (uint32_t i = 0; i < n_grids; ++i)
        if (!grids[i]) return set_error(WX_ERR_INVALID, "batch: NULL grid");
    const bool one = n_reqs == 1u && n_grids == 1u;
    const bool single = one && !(q[0].flags & WX_TRACE_SHADOW);
    if (!one && (n_grids > kMaxGridsLds || n_reqs > kMaxReqsLds))
        return set_error(WX_ERR_INVALID, "batch: at most 64 grids and 1024 requests per launch");
    // device tables of a multi-request launch
    auto tables = [&](const GridDev **g_out, const ReqDev **q_out, const uint32_t **b_out, uint32_t **ctr) -> int {
        if (stg.is_static) {
            *ctr = reinterpret_cast<uint32_t *>(stg.dev + kOffCtr);
            return copy_tables(grids, n_grids, n_reqs, stg.host, stg.dev, stream, g_out, q_out, b_out);
        }
        StreamScratch *S = nullptr;
        int rc = upload_tables(grids, n_grids, q, begin, n_reqs, stream, &S, g_out, q_out, b_out);
        if (rc) return rc;
        *ctr = reinterpret_cast<uint32_t *>(S->dev);
        return WX_OK;
    };
    hipError_t e;
    if (t.persistent == 3u && one && q[0].kind == WX_REQ_CAMERA && grids[0]->dev.n5count == 1u) {
        // LDS shape: the node-5 records in LDS
        dim3 grid((q[0].w + 31u) / 32u, (q[0].h + 31u) / 32u, 1);
        if (q[0].flags & WX_TRACE_SHADOW)
            hipLaunchKernelGGL(k_lds_camera<true>, grid, dim3(1024), 0, stream, grids[0]->dev, q[0]);
        else
            hipLaunchKernelGGL(k_lds_camera<false>, grid, dim3(1024), 0, stream, grids[0]->dev, q[0]);
        e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(e, "batch: launch");
        return WX_OK;
    }
    if (t.persistent == 2u && single && q[0].kind == WX_REQ_CAMERA) {
        // compact shape: primary camera rays, `refill` iterations between censuses
        dim3 grid((q[0].w + 15u) / 16u, (q[0].h + 15u) / 16u, 1);
        hipLaunchKernelGGL(k_compact_camera, grid, dim3(256), 0, stream, grids[0]->dev, q[0], t.refill);
        e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(e, "batch: launch");
        return WX_OK;
    }
    if (t.persistent != 1u) {  // the one-lane shape (also for what the compact / LDS shapes do not cover)
        if (one) {
            const bool sh = (q[0].flags & WX_TRACE_SHADOW) != 0u;
            if (q[0].kind == WX_REQ_CAMERA) {
                dim3 grid((q[0].w + 15u) / 16u, (q[0].h + 7u) / 8u, 1);
                const uint32_t band = t.xcd_split == 2u ? 1u : 0u;  // XCD-banded tile order
                if (sh) hipLaunchKernelGGL(k_simple_camera<true>, grid, dim3(128), 0, stream, grids[0]->dev, q[0], band);
                else hipLaunchKernelGGL(k_simple_camera<false>, grid, dim3(128), 0, stream, grids[0]->dev, q[0], band);
            } else {
                dim3 grid((q[0].w + 127u) / 128u);
                if (sh) hipLaunchKernelGGL(k_simple_rays<true>, grid, dim3(128), 0, stream, grids[0]->dev, q[0]);
                else hipLaunchKernelGGL(k_simple_rays<false>, grid, dim3(128), 0, stream, grids[0]->dev, q[0]);
            }
        } else {
            BatchArgs b;
            std::memset(&b, 0, sizeof b);
            uint32_t *ctr = nullptr;
            int rc = tables(&b.grids, &b.reqs, &b.begin, &ctr);
            if (rc) return rc;
            b.n_grids = n_grids;
            b.n_reqs = n_reqs;
            b.total = (uint32_t)total;
            hipLaunchKernelGGL(k_simple_batch, dim3((uint32_t)((total + 127u) / 128u)), dim3(128), 0, stream, b);
        }
        e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(e, "batch: launch");
        return WX_OK;
    }
    PersistArgs a;
    std::memset(&a, 0, sizeof a);
    a.n_grids = n_grids;
    a.n_reqs = n_reqs;
    a.total = (uint32_t)total;
    a.regions = t.xcd_split ? kRegions : 1u;
    a.refill = t.refill;
    a.chunk = t.chunk;
    const int cus = cu_count();
    const uint32_t blocks = std::max<uint32_t>(1u, (uint32_t)cus * t.waves_per_cu / 4u);
    if (single) {
        if (stg.is_static) {
            a.ctr = reinterpret_cast<uint32_t *>(stg.dev + kOffCtr);
        } else {
            StreamScratch *S = nullptr;
            int rc = scratch_for(stream, 1024, 0, &S);
            if (rc) return rc;
            a.ctr = reinterpret_cast<uint32_t *>(S->dev);
        }
    } else {
        int rc = tables(&a.grids, &a.reqs, &a.begin, &a.ctr);
        if (rc) return rc;
    }
    // the queue counters start at zero: a fill kernel (no memset node, so the launch
    // sequence stays valid inside a captured graph)
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(256), 0, stream, a.ctr, 16u * kRegions);
    if (single) {
        if (q[0].kind == WX_REQ_CAMERA)
            hipLaunchKernelGGL(k_persist1<WX_REQ_CAMERA>, dim3(blocks), dim3(256), 0, stream, a, grids[0]->dev, q[0]);
        else
            hipLaunchKernelGGL(k_persist1<WX_REQ_RAYS>, dim3(blocks), dim3(256), 0, stream, a, grids[0]->dev, q[0]);
    } else {
        hipLaunchKernelGGL(k_persist_batch, dim3(blocks), dim3(256), 0, stream, a);
    }
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "batch: launch");
    return WX_OK;
}

// Every index the kernels derive from a descriptor stays inside its pool: root
// entries name a node-5 below n5; a record that can be walked into has its walk
// bits inside its child bits, and its first child plus its child count is at most
// the child pool's size (the child of slot b is first + popcount(kids below b)).
// Records are checked in parallel on the host; returns the first bad one.
int check_desc(const wx_grid_desc *d, const char *who) {
    for (uint32_t k = 0; k < d->n5; ++k) {
        const int32_t idx = d->root[4u * k + 3u];
        if (idx < 0 || (uint32_t)idx >= d->n5)
            return set_error(WX_ERR_INVALID, std::string(who) + ": root entry " + std::to_string(k) +
                                                 " names node-5 " + std::to_string(idx) + " of " +
                                                 std::to_string(d->n5));
    }
    struct Level {
        const uint32_t *rec;
        uint64_t n;
        uint32_t kid_bits, child_pool;
        const char *name;
    };
    const Level lv[2] = {{d->n5_records, (uint64_t)d->n5 * 4096u, 8u, d->n4, "node-5"},
                         {d->n4_records, (uint64_t)d->n4 * 256u, 16u, d->n3, "node-4"}};
    for (const Level &L : lv) {
        if (!L.n) continue;
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        const uint64_t parts = std::min<uint64_t>(std::min<unsigned>(hw, 16u), (L.n + 65535u) / 65536u);
        std::vector<uint64_t> bad(parts, ~0ull);
        auto scan = [&](uint64_t p) {
            const uint64_t b = L.n * p / parts, e = L.n * (p + 1u) / parts;
            const uint32_t km = (1u << L.kid_bits) - 1u;
            for (uint64_t i = b; i < e; ++i) {
                const uint32_t w = L.rec[4u * i], first = L.rec[4u * i + 1u];
                const uint32_t kids = w & km, walk = (w >> L.kid_bits) & km;
                if (!walk) continue;
                if ((walk & ~kids) || (uint64_t)first + (uint64_t)__builtin_popcount(kids) > L.child_pool) {
                    bad[p] = i;
                    return;
                }
            }
        };
        if (parts <= 1u) {
            scan(0);
        } else {
            std::vector<std::thread> th;
            for (uint64_t p = 0; p < parts; ++p) th.emplace_back(scan, p);
            for (auto &t : th) t.join();
        }
        for (uint64_t b : bad)
            if (b != ~0ull)
                return set_error(WX_ERR_INVALID, std::string(who) + ": " + L.name + " record " +
                                                     std::to_string(b) + " walks into a child outside the pool");
    }
    return WX_OK;
}

int trace_one(wx_grid *g, const wx_request &in, uint32_t flags, hipStream_t stream) {
    wx_grid *gs[1] = {g};
    wx_tune t = current_tune();
    if (flags & WX_TRACE_SIMPLE) t.persistent = 0;
    return launch_batch(gs, 1, &in, 1, t, stream);
}

}  // namespace

namespace wx {

// Checks the descriptor, allocates the grid's four pools and copies the root table,
// the node records and, when d->leaves is set, the leaf words from the host; with
// d->leaves NULL the leaf pool is left for the caller to fill. `who` prefixes errors.
int grid_new(const wx_grid_desc *d, const char *who, wx_grid **out) {
    const std::string w(who);
    if ((d->n5 && !d->root) || (d->n5 && !d->n5_records) || (d->n4 && !d->n4_records))
        return set_error(WX_ERR_INVALID, w + ": NULL pool with a non-zero count");
    // the traversal's fast quotient relies on |p| <= bound <= 2^100 after the bound test
    if (!(d->bound > 0.0f) || !(d->bound <= 0x1p100f))
        return set_error(WX_ERR_INVALID, w + ": bound must be in (0, 2^100]");
    // the kernels address the node pools with 32-bit byte offsets (leaves with 64-bit indices)
    if (d->n5 > (1u << 16) || d->n4 > (1u << 20))
        return set_error(WX_ERR_UNSUPPORTED, w + ": more than 65536 node-5 or 2^20 node-4 nodes");
    if (int rc = check_desc(d, who)) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_error(WX_ERR_DEVICE, w + ": no HIP device");
    wx_grid *g = new (std::nothrow) wx_grid();
    if (!g) return set_error(WX_ERR_BUDGET, w + ": host allocation failed");
    (void)hipGetDevice(&g->device);
    const size_t sz[4] = {(size_t)d->n5 * 16u, (size_t)d->n5 * 4096u * 16u, (size_t)d->n4 * 256u * 16u,
                          (size_t)d->n3 * 64u * 4u};
    const void *src[4] = {d->root, d->n5_records, d->n4_records, d->leaves};
    for (int k = 0; k < 4; ++k) {
        size_t n = sz[k] ? sz[k] : 16;
        hipError_t e = hipMalloc(&g->mem[k], n);
        if (e != hipSuccess) {
            g->mem[k] = nullptr;
            wx_grid_destroy(g);
            return set_error(e == hipErrorOutOfMemory ? WX_ERR_BUDGET : WX_ERR_DEVICE,
                             w + ": hipMalloc: " + hipGetErrorString(e));
        }
        g->bytes += n;
        if (sz[k] && src[k]) {
            e = hipMemcpy(g->mem[k], src[k], sz[k], hipMemcpyHostToDevice);
            if (e != hipSuccess) {
                wx_grid_destroy(g);
                return hip_fail(e, (w + ": hipMemcpy").c_str());
            }
        }
    }
    g->dev.root = static_cast<const int4 *>(g->mem[0]);
    g->dev.n5 = static_cast<const uint4 *>(g->mem[1]);
    g->dev.n4 = static_cast<const uint4 *>(g->mem[2]);
    g->dev.leaf = static_cast<const uint32_t *>(g->mem[3]);
    g->dev.n5count = d->n5;
    g->dev.bound = d->bound;
    g->n5 = d->n5;
    g->n4 = d->n4;
    g->n3 = d->n3;
    g->dev.root0 = d->n5 ? make_int4(d->root[0], d->root[1], d->root[2], d->root[3]) : make_int4(0, 0, 0, 0);
    // WX_WIDE_LEAVES=1 takes the 64-bit leaf addressing at any size (to test it on small grids)
    const char *wide = std::getenv("WX_WIDE_LEAVES");
    g->dev.wide = (d->n3 > (1u << 24) || (wide && wide[0] == '1')) ? 1u : 0u;
    *out = g;
    return WX_OK;
}

void *grid_pool(wx_grid *g, int k) { return g->mem[k]; }

}  // namespace wx

extern "C" {

int wx_device_count(int *count) {
    if (!count) return set_error(WX_ERR_INVALID, "wx_device_count: NULL argument");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    *count = (e == hipSuccess) ? n : 0;
    return WX_OK;
}

int wx_grid_create(const wx_grid_desc *d, wx_grid **out) {
    if (!d || !out) return set_error(WX_ERR_INVALID, "wx_grid_create: NULL argument");
    if (d->n3 && !d->leaves) return set_error(WX_ERR_INVALID, "wx_grid_create: NULL pool with a non-zero count");
    return wx::grid_new(d, "wx_grid_create", out);
}

int wx_vdb_grid_create(const wx_vdb *v, float bound, wx_grid **out) {
    if (!v || !out) return set_error(WX_ERR_INVALID, "wx_vdb_grid_create: NULL argument");
    const wx::HostVdb &t = v->tree;
    if (!t.finalized()) return set_error(WX_ERR_STATE, "wx_vdb_grid_create: tree not finalized");
    wx_grid_desc d;
    d.n5 = t.n5();
    d.n4 = t.n4();
    d.n3 = t.n3();
    d.bound = bound;
    d.root = t.root_table().data();
    d.n5_records = reinterpret_cast<const uint32_t *>(t.n5_records().data());
    d.n4_records = reinterpret_cast<const uint32_t *>(t.n4_records().data());
    d.leaves = nullptr;
    // leaf words a device replay left on this device are copied there; otherwise
    // they come from the host copy (fetched first if it is not on the host yet)
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) return set_error(WX_ERR_DEVICE, "wx_vdb_grid_create: no HIP device");
    const wx::DeviceWords *dw = t.device_words();
    const bool on_device = dw && dw->device() == cur;
    if (!on_device) {
        std::string err;
        if (!t.host_words(err)) return set_error(WX_ERR_DEVICE, "wx_vdb_grid_create: " + err);
        d.leaves = t.leaf_slots();
    }
    wx_grid *g = nullptr;
    if (int rc = wx::grid_new(&d, "wx_vdb_grid_create", &g)) return rc;
    if (on_device && d.n3) {
        // copied on a stream of its own and waited for, so the grid is complete for work
        // on any stream when this returns
        hipStream_t st = nullptr;
        hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipMemcpyAsync(g->mem[3], dw->data(), (size_t)d.n3 * 256u, hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (st) (void)hipStreamDestroy(st);
        if (e != hipSuccess) {
            wx_grid_destroy(g);
            return hip_fail(e, "wx_vdb_grid_create: device copy of the leaf words");
        }
    }
    *out = g;
    return WX_OK;
}

int wx_grid_destroy(wx_grid *g) {
    if (!g) return WX_OK;
    for (int k = 0; k < 4; ++k)
        if (g->mem[k]) (void)hipFree(g->mem[k]);
    delete g;
    return WX_OK;
}

int wx_grid_bytes(const wx_grid *g, uint64_t *out) {
    if (!g || !out) return set_error(WX_ERR_INVALID, "wx_grid_bytes: NULL argument");
    *out = g->bytes;
    return WX_OK;
}

int wx_trace_camera(wx_grid *g, const wx_camera *cam, uint32_t w, uint32_t h, float pixel_offset, uint32_t flags,
                    wx_hit *out, wx_stream stream) {
    if (!g || !cam || !out) return set_error(WX_ERR_INVALID, "wx_trace_camera: NULL argument");
    if (flags & ~(WX_TRACE_CLIP | WX_TRACE_SIMPLE)) return set_error(WX_ERR_INVALID, "wx_trace_camera: bad flags");
    wx_request r;
    std::memset(&r, 0, sizeof r);
    r.kind = WX_REQ_CAMERA;
    r.width = w;
    r.height = h;
    r.flags = flags & WX_TRACE_CLIP;
    r.pixel_offset = pixel_offset;
    r.camera = *cam;
    r.out = out;
    return trace_one(g, r, flags, (hipStream_t)stream);
}

int wx_trace_rays(wx_grid *g, const wx_ray *rays, uint32_t n, uint32_t flags, wx_hit *out, wx_stream stream) {
    if (!g || ((!rays || !out) && n)) return set_error(WX_ERR_INVALID, "wx_trace_rays: NULL argument");
    if (flags & ~(WX_TRACE_CLIP | WX_TRACE_SIMPLE)) return set_error(WX_ERR_INVALID, "wx_trace_rays: bad flags");
    wx_request r;
    std::memset(&r, 0, sizeof r);
    r.kind = WX_REQ_RAYS;
    r.width = n;
    r.height = 1;
    r.flags = flags & WX_TRACE_CLIP;
    r.rays = rays;
    r.out = out;
    return trace_one(g, r, flags, (hipStream_t)stream);
}

int wx_tune_get(wx_tune *out) {
    if (!out) return set_error(WX_ERR_INVALID, "wx_tune_get: NULL argument");
    *out = current_tune();
    return WX_OK;
}

int wx_tune_set(const wx_tune *t) {
    if (!t || !tune_ok(*t)) return set_error(WX_ERR_INVALID, "wx_tune_set: out of range");
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[dev] = *t;
    return WX_OK;
}

int wx_trace_batch(wx_grid *const *grids, uint32_t n_grids, const wx_request *reqs, uint32_t n_reqs,
                   const wx_tune *tune, wx_stream stream) {
    if ((!grids && n_grids) || (!reqs && n_reqs)) return set_error(WX_ERR_INVALID, "wx_trace_batch: NULL argument");
    if (n_reqs == 0) return WX_OK;
    if (n_grids == 0) return set_error(WX_ERR_INVALID, "wx_trace_batch: no grids");
    return launch_batch(grids, n_grids, reqs, n_reqs, tune ? *tune : current_tune(), (hipStream_t)stream);
}

int wx_trace_batch_static(wx_grid *const *grids, uint32_t n_grids, const wx_request *reqs, uint32_t n_reqs,
                          const wx_tune *tune, wx_stream stream) {
    if ((!grids && n_grids) || (!reqs && n_reqs))
        return set_error(WX_ERR_INVALID, "wx_trace_batch_static: NULL argument");
    if (n_reqs == 0) return WX_OK;
    if (n_grids == 0) return set_error(WX_ERR_INVALID, "wx_trace_batch_static: no grids");
    if (n_grids > kMaxGridsLds || n_reqs > kMaxReqsLds)
        return set_error(WX_ERR_INVALID, "wx_trace_batch_static: at most 64 grids and 1024 requests per call");
    Staging stg;
    stg.is_static = true;
    if (int rc = static_scratch(&stg.dev, &stg.host)) return rc;
    return launch_batch(grids, n_grids, reqs, n_reqs, tune ? *tune : current_tune(), (hipStream_t)stream, stg);
}

int wx_autotune(wx_grid *const *grids, uint32_t n_grids, const wx_request *reqs, uint32_t n_reqs, wx_stream stream,
                wx_tune *best, float *best_ms) {
    if (!grids || !reqs || !n_grids || !n_reqs) return set_error(WX_ERR_INVALID, "wx_autotune: empty batch");
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return set_error(WX_ERR_DEVICE, "wx_autotune: event create");
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return set_error(WX_ERR_DEVICE, "wx_autotune: event create");
    }
    const uint32_t refills[] = {1, 8, 16, 24, 32, 48};
    const uint32_t chunks[] = {64, 128, 256, 512};
    const uint32_t waves[] = {8, 12, 16, 24, 32};
    wx_tune bt = default_tune();
    float bms = 1e30f;
    int rc = WX_OK;
    auto time_it = [&](const wx_tune &t, float &ms) -> int {
        int r = launch_batch(grids, n_grids, reqs, n_reqs, t, s);  // warm
        if (r) return r;
        float acc = 0.f;
        for (int rep = 0; rep < 3; ++rep) {
            if (hipEventRecord(e0, s) != hipSuccess) return set_error(WX_ERR_DEVICE, "wx_autotune: event record");
            r = launch_batch(grids, n_grids, reqs, n_reqs, t, s);
            if (r) return r;
            if (hipEventRecord(e1, s) != hipSuccess) return set_error(WX_ERR_DEVICE, "wx_autotune: event record");
            if (hipEventSynchronize(e1) != hipSuccess) return set_error(WX_ERR_DEVICE, "wx_autotune: sync");
            float m = 0.f;
            if (hipEventElapsedTime(&m, e0, e1) != hipSuccess) return set_error(WX_ERR_DEVICE, "wx_autotune: event time");
            acc += m;
        }
        ms = acc / 3.f;
        return WX_OK;
    };
    // the one-lane-per-ray shape, the LDS shape for a lone camera request on a grid
    // with one node-5, the compact shape for a lone primary camera request (census
    // intervals), then a coordinate search over the persistent shape: waves, refill,
    // chunk, XCD split
    {
        wx_tune c = default_tune();
        c.persistent = 0;
        float ms;
        rc = time_it(c, ms);
        if (rc == WX_OK && ms < bms) bms = ms, bt = c;
    }
    if (rc == WX_OK && n_reqs == 1u && n_grids == 1u && reqs[0].kind == WX_REQ_CAMERA &&
        grids[0]->dev.n5count == 1u) {
        wx_tune c = default_tune();
        c.persistent = 3;
        float ms;
        rc = time_it(c, ms);
        if (rc == WX_OK && ms < bms) bms = ms, bt = c;
    }
    if (rc == WX_OK && n_reqs == 1u && n_grids == 1u && reqs[0].kind == WX_REQ_CAMERA &&
        !(reqs[0].flags & WX_TRACE_SHADOW)) {
        for (uint32_t k : {4u, 8u, 16u}) {
            wx_tune c = default_tune();
            c.persistent = 2;
            c.refill = k;
            float ms;
            if ((rc = time_it(c, ms))) break;
            if (ms < bms) bms = ms, bt = c;
        }
    }
    wx_tune t = default_tune();
    t.persistent = 1;
    wx_tune pbest = t;
    float pms = 1e30f;
    for (int pass = 0; pass < 2 && rc == WX_OK; ++pass) {
        for (uint32_t w : waves) {
            wx_tune c = t;
            c.waves_per_cu = w;
            float ms;
            if ((rc = time_it(c, ms))) break;
            if (ms < pms) pms = ms, pbest = c;
        }
        t = pbest;
        for (uint32_t rf : refills) {
            wx_tune c = t;
            c.refill = rf;
            float ms;
            if ((rc = time_it(c, ms))) break;
            if (ms < pms) pms = ms, pbest = c;
        }
        t = pbest;
        for (uint32_t ch : chunks) {
            wx_tune c = t;
            c.chunk = ch;
            float ms;
            if ((rc = time_it(c, ms))) break;
            if (ms < pms) pms = ms, pbest = c;
        }
        t = pbest;
        for (uint32_t x = 0; x <= 1; ++x) {
            wx_tune c = t;
            c.xcd_split = x;
            float ms;
            if ((rc = time_it(c, ms))) break;
            if (ms < pms) pms = ms, pbest = c;
        }
        t = pbest;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (pms < bms) bms = pms, bt = pbest;
    wx_tune_set(&bt);
    if (best) *best = bt;
    if (best_ms) *best_ms = bms;
    return WX_OK;
}

}  // extern "C"
