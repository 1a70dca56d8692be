// Serving loop: resident grids, HBM-budget admission and continuous batching.
//
// Requests are queued FIFO. A dispatcher thread owns `streams` launch slots (one
// HIP stream each, default priority): whenever a slot is free it packs queued
// requests whose outputs fit in the remaining budget into ONE batched launch
// (wx_trace_batch), so requests that arrive while other launches run on the GPU
// join the next launch and overlap with them on the device. One completion
// thread per stream waits for its launches in order and timestamps the requests.
// Output buffers come from one device arena allocated at creation and
// sub-allocated first-fit, so serving never calls hipMalloc.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "lane.hpp"
#include "request.hpp"
#include "wx_internal.hpp"

using wx::set_error;

namespace {

using namespace wxk;

// ---------------------------------------------------------------- resident sessions
//
// One kernel (k_serve) runs for a session while the host admits requests into it.
// Each request's 8x8 tiles are dealt in pairs over kXcdSeqs bands (pair k to band
// k % kXcdSeqs, as the batched kernels' workgroups of two tiles fall to the XCDs), band
// x joining tile sequence x; workgroup b works on sequence b % kXcdSeqs, the XCD the
// hardware dispatches it to, so each claim counter serves one XCD and every XCD has
// an even share of every request. The host writes each request and its band ends into a
// mailbox in pinned host memory, then raises the request count. One wave of the kernel
// polls the mailbox, copies new requests into device memory and publishes them there;
// the other waves claim tiles in chunks from their sequence's device counter, wait
// (bounded) until a claimed tile is published, trace it one lane per pixel (the
// one-request kernel's loop, shadow ray continued in the lane), write the hit records
// through to memory and count the tile against its request. The wave that completes
// a request flags it in the mailbox. Every wait ends on the host's stop or on the
// session's wait limit.

constexpr uint32_t kSessionReqs = 16384u;  // requests one session admits
constexpr uint32_t kSessionGrids = 64u;    // grids one session can trace
constexpr uint32_t kSessionChunk = 8u;     // most tiles per claim (measured: 4 and 16 slower)
constexpr uint32_t kSessionMaxTiles = 0x7FFF0000u;  // per sequence
constexpr uint32_t kXcdSeqs = 8u;                   // tile sequences, one per XCD

struct SessionReq {  // one admitted request (mailbox)
    ReqDev q;
    uint32_t end[kXcdSeqs];  // where its band ends in each sequence
    uint32_t tiles;          // its tile count
    uint32_t pad[3];
};
static_assert(sizeof(SessionReq) == 176, "SessionReq layout");

constexpr uint32_t kBandPair = 2u;  // consecutive tiles dealt to one band

// Tiles of band x of a request of T tiles: those of the pairs k = x, x + 8, ...
__host__ __device__ __forceinline__ uint32_t band_size(uint32_t T, uint32_t x) {
    const uint32_t round = kBandPair * kXcdSeqs, r = T % round, lo = kBandPair * x;
    return kBandPair * (T / round) + (r > lo ? (r - lo < kBandPair ? r - lo : kBandPair) : 0u);
}
// The request tile at position p of band x.
__host__ __device__ __forceinline__ uint32_t band_tile(uint32_t p, uint32_t x) {
    return (p / kBandPair) * (kBandPair * kXcdSeqs) + kBandPair * x + p % kBandPair;
}

struct Mailbox {  // pinned host memory, read and written by both sides with coherent accesses
    uint32_t published_reqs;  // host: admitted requests (raised after their slots)
    uint32_t stop;            // host: nothing more is admitted; waves end once out of work
    uint32_t timed_out;       // kernel: a wait ended on the wait limit
    uint32_t pad[29];
    uint32_t done[kSessionReqs];  // kernel: every hit record of request k is in memory
    SessionReq ring[kSessionReqs];
};

struct SessionDev {  // device memory, zeroed at session start
    uint32_t tiles[kXcdSeqs], reqs, stop;  // the poller's mirror of the mailbox counts
    uint32_t pad0[22];
    struct {
        uint32_t v, pad[31];
    } next[kXcdSeqs];  // claim counter of each sequence, a line each
    uint32_t done_tiles[kSessionReqs];
    GridDev grids[kSessionGrids];
    // the poller's copies of the admitted requests, written before they are published
    // (so the waves read them from device memory, not across PCIe)
    uint32_t ends[kXcdSeqs][kSessionReqs];  // SessionReq::end, by sequence
    uint32_t recs[kSessionReqs][64];        // the ReqDev words, then the tile count (256 B each)
};

struct ServeArgs {
    Mailbox *mb;  // device address of the mailbox
    SessionDev *sd;
    const char *arena;  // outputs must lie in [arena, arena + arena_bytes)
    uint64_t arena_bytes;
    uint32_t n_grids, chunk;
    uint64_t wait_ticks;     // wait limit in s_memrealtime ticks (100 MHz)
    uint32_t waves_per_seq;  // worker waves of each sequence
    uint32_t pad;
};

// A request record the kernel may trace: a known grid and outputs inside the arena.
__device__ __forceinline__ bool request_ok(const ServeArgs &a, const ReqDev &q, uint32_t tiles) {
    const uint64_t px = (uint64_t)q.w * q.h * 32u, lo = (uint64_t)a.arena, hi = lo + a.arena_bytes;
    const uint64_t o = (uint64_t)q.out, sh = (uint64_t)q.shadow;
    return q.grid < a.n_grids && q.tiles_x == (q.w + 7u) / 8u &&
           (uint64_t)tiles == (uint64_t)q.tiles_x * ((q.h + 7u) / 8u) && o >= lo && o + px <= hi &&
           (!(q.flags & WX_TRACE_SHADOW) || (sh >= lo && sh + px <= hi));
}

__device__ __forceinline__ uint32_t sys_load(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ void sys_store(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ uint32_t dev_load(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void dev_store(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
typedef __attribute__((address_space(1))) uint32_t gu32;
__device__ __forceinline__ void dev_store(gu32 *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Waits until every memory access this wave issued has completed: device-scope
// stores issued before are then coherent in memory for every other agent.
__device__ __forceinline__ void wait_memory() { __builtin_amdgcn_s_waitcnt(0); }
__device__ __forceinline__ uint64_t ticks() { return __builtin_amdgcn_s_memrealtime(); }

// A hit record written with device-scope stores: coherent in memory, not held in
// this XCD's L2, once the wave's stores have completed.
__device__ __forceinline__ void write_through(const RayLane &r, float ox, float oy, float oz, wx_hit *out) {
    uint4 w0, w1;
    hit_words(r, ox, oy, oz, w0, w1);
    gu32 *o = (gu32 *)reinterpret_cast<uint32_t *>(out);  // the output arena is global memory
    dev_store(o + 0, w0.x), dev_store(o + 1, w0.y), dev_store(o + 2, w0.z), dev_store(o + 3, w0.w);
    dev_store(o + 4, w1.x), dev_store(o + 5, w1.y), dev_store(o + 6, w1.z), dev_store(o + 7, w1.w);
}
__device__ __forceinline__ void write_through_no_ray(wx_hit *out) {
    gu32 *o = (gu32 *)reinterpret_cast<uint32_t *>(out);
    for (int k = 0; k < 8; ++k) dev_store(o + k, k == 6 ? (uint32_t)WX_NO_RAY : 0u);
}

// The poller: wave 0 of workgroup 0. Every ~1 us it copies newly admitted requests
// (up to 16 at a time) into device memory, then publishes them there: the request
// count, then each sequence's length (the band end of the last copied request),
// each store completed before the next. It mirrors stop once every admitted request
// is published, and ends then or when nothing new was admitted for the wait limit.
__device__ void serve_poll(const ServeArgs &a) {
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t copied = 0, ends = 0;  // lane 32 + x holds sequence x's published length
    uint64_t t_last = ticks();
    for (;;) {
        // stop before the request count: a stop seen covers every admission
        const uint32_t stop = (uint32_t)__builtin_amdgcn_readfirstlane((int)sys_load(&a.mb->stop));
        wait_memory();
        const uint32_t reqs = (uint32_t)__builtin_amdgcn_readfirstlane((int)sys_load(&a.mb->published_reqs));
        wait_memory();
        const uint32_t target = reqs < kSessionReqs ? reqs : kSessionReqs;
        uint32_t n = 0;
        for (; copied < target && n < 16u; ++copied, ++n) {  // 41 words, one per lane
            const uint32_t *src = reinterpret_cast<const uint32_t *>(&a.mb->ring[copied]);
            const uint32_t v = lane < 41u ? sys_load(src + lane) : 0u;
            if (lane < 32u) {
                dev_store(&a.sd->recs[copied][lane], v);
            } else if (lane < 32u + kXcdSeqs) {
                dev_store(&a.sd->ends[lane - 32u][copied], v);
                ends = v;
            } else if (lane == 40u) {
                dev_store(&a.sd->recs[copied][32], v);
            }
        }
        const uint64_t t = ticks();
        if (n) {
            wait_memory();
            if (lane == 0) dev_store(&a.sd->reqs, copied);
            wait_memory();
            if (lane >= 32u && lane < 32u + kXcdSeqs) dev_store(&a.sd->tiles[lane - 32u], ends);
            wait_memory();
            t_last = t;
            continue;  // more may be waiting: no sleep
        }
        if (stop) {
            if (lane == 0) dev_store(&a.sd->stop, 1u);
            return;
        }
        if (t - t_last > a.wait_ticks) {
            sys_store(&a.mb->timed_out, 1u);
            dev_store(&a.sd->stop, 1u);
            return;
        }
        __builtin_amdgcn_s_sleep(32);
    }
}

// 8 waves per SIMD: the persistent loop keeps loop-invariant values live, which
// would otherwise hold it to 5 (measured 16 % slower on a burst of 256 requests);
// the few spills it costs sit outside the traversal loops.
__global__ __launch_bounds__(256, 8) void k_serve(ServeArgs a) {
    __shared__ float2 s_tab[kTab];
    init_tab(s_tab);
    if (blockIdx.x == 0 && threadIdx.x < 64u) {
        serve_poll(a);
        return;
    }
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t x = blockIdx.x % kXcdSeqs;  // this workgroup's sequence (and XCD)
    uint32_t cur = 0, end = 0, known = 0;       // positions in sequence x
    uint32_t req = 0, req_begin = 0, req_end = 0, req_tiles = 0;
    uint32_t counted = 0;  // tiles of `req` traced by this wave and not yet counted
    bool have = false;
    ReqDev q;
    // Counts this wave's traced tiles against their request once their hit records are
    // in memory; the wave that counts a request's last tile flags it in the mailbox.
    auto count = [&]() {
        if (!counted) return;
        wait_memory();
        if (lane == 0) {
            const uint32_t old = __hip_atomic_fetch_add(&a.sd->done_tiles[req], counted, __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT);
            if (old + counted == req_tiles) sys_store(&a.mb->done[req], 1u);
        }
        counted = 0;
    };
    for (;;) {
        if (cur == end) {  // count the chunk's tiles, claim the next chunk
            count();
            // the claim is sized by the sequence's backlog (published, not yet claimed, per
            // wave), so a request that arrives at low load spreads over every wave of the
            // sequence instead of queueing in a few waves' chunks
            known = (uint32_t)__builtin_amdgcn_readfirstlane((int)dev_load(&a.sd->tiles[x]));
            const uint32_t nx = (uint32_t)__builtin_amdgcn_readfirstlane((int)dev_load(&a.sd->next[x].v));
            const uint32_t backlog = known > nx ? (known - nx) / a.waves_per_seq : 0u;
            const uint32_t c = backlog < 1u ? 1u : (backlog < a.chunk ? backlog : a.chunk);
            uint32_t got = 0;
            if (lane == 0) got = atomicAdd(&a.sd->next[x].v, c);
            cur = (uint32_t)__builtin_amdgcn_readfirstlane((int)got);
            end = cur + c;
        }
        if (cur >= known) {  // wait until tile `cur` is published (bounded)
            count();
            const uint64_t t0 = ticks();
            uint32_t backoff = 1;
            bool quit = false;
            for (;;) {  // every value here is wave-uniform
                known = (uint32_t)__builtin_amdgcn_readfirstlane((int)dev_load(&a.sd->tiles[x]));
                if (cur < known) break;
                if (__builtin_amdgcn_readfirstlane((int)dev_load(&a.sd->stop))) {
                    // stop is mirrored after the final count: read the count again
                    known = (uint32_t)__builtin_amdgcn_readfirstlane((int)dev_load(&a.sd->tiles[x]));
                    quit = cur >= known;
                    break;
                }
                if (ticks() - t0 > a.wait_ticks) {
                    if (lane == 0) sys_store(&a.mb->timed_out, 1u);
                    quit = true;
                    break;
                }
                for (uint32_t k = 0; k < backoff; ++k) __builtin_amdgcn_s_sleep(32);
                backoff = backoff < 64u ? backoff * 2u : 64u;
            }
            if (quit) break;
        }
        if (!have || cur >= req_end) {
            count();
            // the request whose band holds position `cur`, in [lo, hi): requests are
            // admitted in order, so their band ends in a sequence do not decrease; a
            // 64-way search, each lane probing one request (the requests that end at or
            // before `cur` form a prefix)
            const uint32_t R = (uint32_t)__builtin_amdgcn_readfirstlane((int)dev_load(&a.sd->reqs));
            uint32_t lo = have ? req + 1u : 0u, hi = R;
            while (hi > lo + 1u) {
                const uint32_t stride = (hi - lo + 63u) / 64u;
                const uint32_t j = lo + lane * stride;
                const bool past = j < hi && cur >= dev_load(&a.sd->ends[x][j]);
                const uint32_t n = (uint32_t)__popcll(__ballot(past));
                if (n == 0u) break;  // request lo holds it
                const uint32_t next_hi = lo + n * stride + 1u;  // probe n does not end before `cur`
                lo += (n - 1u) * stride + 1u;
                hi = hi < next_hi ? hi : next_hi;
            }
            // tile `cur` is published, so is its request (lo < R); the clamp only keeps the
            // read inside the ring, and a record not holding `cur` is skipped below
            req = lo < R ? lo : (R ? R - 1u : 0u);
            // the request record (32 words and the tile count) and its band end, one word
            // per lane, gathered into scalars
            const uint32_t v = lane < 33u ? dev_load(&a.sd->recs[req][lane])
                                          : (lane == 33u ? dev_load(&a.sd->ends[x][req]) : 0u);
            uint32_t w[32];
#pragma unroll
            for (int k = 0; k < 32; ++k) w[k] = (uint32_t)__builtin_amdgcn_readlane((int)v, k);
            __builtin_memcpy(&q, w, sizeof q);
            const uint32_t T = (uint32_t)__builtin_amdgcn_readlane((int)v, 32);
            req_tiles = T;
            req_end = (uint32_t)__builtin_amdgcn_readlane((int)v, 33);
            req_begin = req_end - band_size(T, x);
            have = true;
        }
        const bool ok = request_ok(a, q, req_tiles) && cur >= req_begin && cur < req_end;
        if (ok) {
            const GridDev g = a.sd->grids[q.grid];
            const uint32_t slot = band_tile(cur - req_begin, x) * 64u + lane;
            uint32_t px, py;
            if (camera_pixel(q, slot, px, py)) {  // lanes outside the image sit the tile out
                const uint32_t pix = py * q.w + px;
                float dx, dy, dz;
                camera_dir(q, px, py, dx, dy, dz);
                RayLane r;
                lane_start(r, q.ex, q.ey, q.ez, dx, dy, dz, g.bound, (q.flags & WX_TRACE_CLIP) != 0u);
                lane_run(g, r, s_tab);
                write_through(r, q.ex, q.ey, q.ez, q.out + pix);
                if (q.flags & WX_TRACE_SHADOW) {
                    if (lane_state(r) == WX_HIT) {
                        float ox, oy, oz;
                        shadow_origin(r, ox, oy, oz);
                        start_shadow(r, q, g.bound, ox, oy, oz);
                        lane_run(g, r, s_tab);
                        write_through(r, ox, oy, oz, q.shadow + pix);
                    } else {
                        write_through_no_ray(q.shadow + pix);
                    }
                }
            }
        }
        if (ok) ++counted;  // a record that fails the checks is never counted done
        ++cur;
    }
}

__global__ __launch_bounds__(256) void k_fill_zero(uint32_t *p, uint32_t n) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) p[i] = 0u;
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// First-fit allocator over [0, cap) in 256-byte units.
class Arena {
  public:
    void init(uint64_t cap) {
        free_.clear();
        if (cap) free_[0] = cap;
    }
    static uint64_t round(uint64_t n) { return (n + 255u) & ~uint64_t(255); }
    bool alloc(uint64_t n, uint64_t *off) {
        n = round(n);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            *off = it->first;
            const uint64_t rest = it->second - n, start = it->first + n;
            free_.erase(it);
            if (rest) free_[start] = rest;
            return true;
        }
        return false;
    }
    void release(uint64_t off, uint64_t n) {
        n = round(n);
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }

  private:
    std::map<uint64_t, uint64_t> free_;
};

enum JobState { Q_QUEUED, Q_LAUNCHED, Q_DONE, Q_CANCELLED };

struct Job {
    wx_serve_request req;
    uint64_t bytes = 0, off = 0;
    JobState state = Q_QUEUED;
    double t_submit = 0, t_launch = 0, t_done = 0;
    uint32_t batch_n = 0;
    int status = WX_OK;
};

struct Batch {
    std::vector<uint64_t> ids;
    hipEvent_t ev = nullptr;
    int status = WX_OK;
};

struct Slot {
    hipStream_t stream = nullptr;
    std::deque<Batch> inflight;  // at most one launched batch at a time
};

constexpr size_t kMaxBatchGrids = 64;               // grids one batched launch takes
constexpr uint64_t kMaxBatchSlots = 0x7FFFFFFFull;  // ray slots one batched launch takes

uint64_t job_bytes(const wx_serve_request &r) {
    const uint64_t px = (uint64_t)r.width * r.height;
    return px * 32u * ((r.flags & WX_TRACE_SHADOW) ? 2u : 1u);
}

// Latencies in log-spaced bins: bin k (1 <= k < kBins-1) holds [kLo * r^(k-1),
// kLo * r^k) ms with r = 1.01; bin 0 everything below kLo, the last bin everything
// above. Fixed size, so the record never grows; a percentile is read to within
// one bin (1 %).
constexpr double kLatLoMs = 1e-3, kLatRatio = 1.01;
constexpr size_t kLatBins = 2086;  // up to ~1e6 ms

struct LatencyHist {
    uint64_t count[kLatBins] = {};
    uint64_t n = 0;
    double sum = 0.0;

    void add(double ms) {
        size_t k = 0;
        if (ms >= kLatLoMs) k = std::min(kLatBins - 1, (size_t)(std::log(ms / kLatLoMs) / std::log(kLatRatio)) + 1u);
        count[k]++;
        n++;
        sum += ms;
    }
    // The geometric centre of the bin holding the q-quantile.
    double percentile(double q) const {
        if (!n) return 0.0;
        const uint64_t target = std::min<uint64_t>(n - 1u, (uint64_t)(q * (double)(n - 1u)));
        uint64_t seen = 0;
        for (size_t k = 0; k < kLatBins; ++k) {
            seen += count[k];
            if (seen > target) {
                if (k == 0) return kLatLoMs;
                return kLatLoMs * std::pow(kLatRatio, (double)k - 0.5);
            }
        }
        return kLatLoMs * std::pow(kLatRatio, (double)kLatBins);
    }
};

}  // namespace

struct wx_server {
    wx_server_config cfg{};
    bool dry = false;
    int device = 0;
    std::vector<wx_grid *> grids;
    char *arena_dev = nullptr;
    Arena arena;
    uint64_t in_use = 0, peak = 0;
    std::vector<Slot> slots;

    std::mutex mu;
    std::condition_variable cv_dispatch, cv_done, cv_slot;
    std::deque<uint64_t> queue;
    std::unordered_map<uint64_t, Job> jobs;
    uint64_t next_id = 1;
    bool stop = false;

    wx_server_stats st{};
    LatencyHist lat;
    double batch_req_sum = 0;

    std::thread dispatcher;
    std::vector<std::thread> completers;

    // resident sessions (WX_SERVER_RESIDENT)
    bool resident = false;
    Mailbox *mb = nullptr;      // pinned host memory
    Mailbox *mb_dev = nullptr;  // its device address
    SessionDev *sd = nullptr;
    hipStream_t s_stream = nullptr;
    hipEvent_t s_end = nullptr;
    int s_blocks = 0;
    bool s_running = false, s_stopping = false;
    uint32_t s_grids = 0, s_reqs = 0, s_done = 0, s_fruitless = 0;
    uint32_t s_seq[kXcdSeqs] = {};  // published length of each tile sequence
    std::vector<uint64_t> s_ids;      // job of each admitted request, by session index
    std::vector<uint32_t> s_pending;  // admitted, not yet seen done
    double s_last_activity = 0;
    std::condition_variable cv_session;

    void dispatch_loop();
    void complete_loop(size_t k);
    void resident_dispatch_loop();
    void resident_complete_loop();
    int start_session();
    void end_session();
    void fail_job(uint64_t id, int status, double t);
};

void wx_server::fail_job(uint64_t id, int status, double t) {
    Job &j = jobs[id];
    j.state = Q_DONE;
    j.t_done = t;
    j.status = status;
    st.completed++;
}

// Under the lock. Snapshot of the grids, device state zeroed, mailbox reset, kernel launched.
int wx_server::start_session() {
    s_grids = (uint32_t)std::min<size_t>(grids.size(), kSessionGrids);
    std::vector<GridDev> table(s_grids);
    for (uint32_t i = 0; i < s_grids; ++i) table[i] = grids[i]->dev;
    std::memset(mb, 0, offsetof(Mailbox, ring));
    const uint32_t words = (uint32_t)(offsetof(SessionDev, grids) / 4u);
    // no kernel of an earlier session is running: its end event has completed
    if (s_grids && hipMemcpy(sd->grids, table.data(), s_grids * sizeof(GridDev), hipMemcpyHostToDevice) != hipSuccess)
        return set_error(WX_ERR_DEVICE, "server: session grid table upload");
    hipLaunchKernelGGL(k_fill_zero, dim3((words + 255u) / 256u), dim3(256), 0, s_stream,
                       reinterpret_cast<uint32_t *>(sd), words);
    ServeArgs a;
    a.mb = mb_dev;
    a.sd = sd;
    a.arena = arena_dev;
    a.arena_bytes = cfg.hbm_budget;
    a.n_grids = s_grids;
    a.chunk = kSessionChunk;
    a.wait_ticks = (uint64_t)(cfg.wait_limit_us ? cfg.wait_limit_us : 2000000u) * 100u;  // 100 MHz clock
    a.waves_per_seq = std::max(1u, (uint32_t)s_blocks * 4u / kXcdSeqs);
    a.pad = 0;
    hipLaunchKernelGGL(k_serve, dim3((uint32_t)s_blocks), dim3(256), 0, s_stream, a);
    if (hipGetLastError() != hipSuccess || hipEventRecord(s_end, s_stream) != hipSuccess)
        return set_error(WX_ERR_DEVICE, "server: session launch");
    s_running = true;
    s_stopping = false;
    s_reqs = s_done = 0;
    std::fill(s_seq, s_seq + kXcdSeqs, 0u);
    s_ids.clear();
    s_pending.clear();
    s_last_activity = now_s();
    st.batches++;
    cv_session.notify_all();
    return WX_OK;
}

// Under the lock: nothing more is admitted; the kernel ends once its waves are out of
// published work.
void wx_server::end_session() {
    s_stopping = true;
    __atomic_store_n(&mb->stop, 1u, __ATOMIC_RELEASE);
}

void wx_server::resident_dispatch_loop() {
    (void)hipSetDevice(device);
    const double idle_s = (cfg.idle_stop_us ? cfg.idle_stop_us : 50000u) * 1e-6;
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        if (stop) {
            for (uint64_t id : queue) {
                jobs[id].state = Q_CANCELLED;
                st.cancelled++;
            }
            queue.clear();
            if (s_running && !s_stopping) end_session();
            cv_done.notify_all();
            cv_session.notify_all();
            break;
        }
        if (!s_running && !queue.empty()) {
            if (int rc = start_session()) {  // the head request fails; the rest wait for another try
                const uint64_t id = queue.front();
                queue.pop_front();
                fail_job(id, rc, now_s());
                cv_done.notify_all();
                continue;
            }
        }
        while (s_running && !s_stopping && !queue.empty()) {
            const uint64_t id = queue.front();
            Job &j = jobs[id];
            const uint32_t tx = (j.req.width + 7u) / 8u, ty = (j.req.height + 7u) / 8u;
            const uint32_t tiles = tx * ty;  // at most 2^26
            bool full = false;
            for (uint32_t x = 0; x < kXcdSeqs; ++x)
                full |= (uint64_t)s_seq[x] + band_size(tiles, x) > kSessionMaxTiles;
            if (j.req.grid >= s_grids || s_reqs == kSessionReqs || full) {
                end_session();  // the next session takes it
                break;
            }
            uint64_t off = 0;
            if (!arena.alloc(j.bytes, &off)) break;  // wait for memory (keeps FIFO order)
            j.off = off;
            in_use += Arena::round(j.bytes);
            peak = std::max(peak, in_use);
            const uint32_t k = s_reqs;
            SessionReq &slot = mb->ring[k];
            ReqDev &q = slot.q;
            std::memset(&q, 0, sizeof q);
            q.kind = WX_REQ_CAMERA;
            q.grid = j.req.grid;
            q.w = j.req.width, q.h = j.req.height;
            q.flags = j.req.flags;
            q.tiles_x = tx;
            q.off = j.req.pixel_offset;
            const wx_camera &c = j.req.camera;
            q.ex = c.eye[0], q.ey = c.eye[1], q.ez = c.eye[2];
            q.ux = c.u[0], q.uy = c.u[1], q.uz = c.u[2];
            q.mx = c.mv[0], q.my = c.mv[1], q.mz = c.mv[2];
            q.wx = c.wp[0], q.wy = c.wp[1], q.wz = c.wp[2];
            q.lx = j.req.light[0], q.ly = j.req.light[1], q.lz = j.req.light[2];
            const uint64_t px = (uint64_t)j.req.width * j.req.height;
            q.out = reinterpret_cast<wx_hit *>(arena_dev + off);
            q.shadow = (j.req.flags & WX_TRACE_SHADOW) ? reinterpret_cast<wx_hit *>(arena_dev + off + px * 32u) : nullptr;
            for (uint32_t x = 0; x < kXcdSeqs; ++x) {  // band x joins sequence x
                s_seq[x] += band_size(tiles, x);
                slot.end[x] = s_seq[x];
            }
            slot.tiles = tiles;
            // the slot, then the request count that publishes it
            __atomic_store_n(&mb->published_reqs, k + 1u, __ATOMIC_RELEASE);
            s_reqs = k + 1u;
            s_ids.push_back(id);
            s_pending.push_back(k);
            queue.pop_front();
            j.state = Q_LAUNCHED;
            j.t_launch = now_s();
            j.batch_n = (uint32_t)s_pending.size();
            batch_req_sum += 1.0;
            s_last_activity = j.t_launch;
        }
        cv_session.notify_all();
        if (s_running && !s_stopping && queue.empty() && s_pending.empty()) {
            const double idle = now_s() - s_last_activity;
            if (idle >= idle_s) {
                end_session();
                continue;
            }
            cv_dispatch.wait_for(lk, std::chrono::duration<double>(idle_s - idle));
            continue;
        }
        cv_dispatch.wait(lk);
    }
}

void wx_server::resident_complete_loop() {
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        if (!s_running) {
            if (stop) break;
            cv_session.wait(lk);
            continue;
        }
        const std::vector<uint32_t> pend = s_pending;
        lk.unlock();
        std::vector<uint32_t> fin;
        for (uint32_t k : pend)
            if (__atomic_load_n(&mb->done[k], __ATOMIC_ACQUIRE)) fin.push_back(k);
        bool ended = false;
        if (fin.size() < pend.size() || pend.empty()) {
            ended = hipEventQuery(s_end) == hipSuccess;
            if (ended) {  // flags raised before the kernel ended
                fin.clear();
                for (uint32_t k : pend)
                    if (__atomic_load_n(&mb->done[k], __ATOMIC_ACQUIRE)) fin.push_back(k);
            }
        }
        const double t = now_s();
        lk.lock();
        for (uint32_t k : fin) {
            Job &j = jobs[s_ids[k]];
            j.state = Q_DONE;
            j.t_done = t;
            j.status = WX_OK;
            st.completed++;
            st.primary_rays += (uint64_t)j.req.width * j.req.height;
            lat.add((t - j.t_submit) * 1e3);
            s_pending.erase(std::find(s_pending.begin(), s_pending.end(), k));
            s_done++;
        }
        if (!fin.empty()) s_last_activity = t;
        if (ended) {
            // requests the session left unfinished (a wait ended on its limit) run again
            // in the next session; after three sessions that finish nothing they fail
            s_fruitless = (s_done == 0 && !s_pending.empty()) ? s_fruitless + 1u : 0u;
            for (auto it = s_pending.rbegin(); it != s_pending.rend(); ++it) {
                const uint64_t id = s_ids[*it];
                Job &j = jobs[id];
                arena.release(j.off, j.bytes);
                in_use -= Arena::round(j.bytes);
                if (stop) {  // shutting down: not run again
                    j.state = Q_CANCELLED;
                    st.cancelled++;
                } else if (s_fruitless >= 3u) {
                    fail_job(id, set_error(WX_ERR_DEVICE, "server: the session kernel made no progress"), t);
                } else {
                    j.state = Q_QUEUED;
                    queue.push_front(id);
                }
            }
            s_pending.clear();
            s_running = false;
            s_stopping = false;
        }
        if (!fin.empty() || ended) {
            cv_done.notify_all();
            cv_dispatch.notify_all();
            continue;
        }
        // nothing new: poll again shortly
        lk.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(10));
        lk.lock();
    }
}

void wx_server::dispatch_loop() {
    if (!dry) (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        if (stop) {  // shutting down: requests not yet launched are cancelled
            for (uint64_t id : queue) {
                jobs[id].state = Q_CANCELLED;
                st.cancelled++;
            }
            queue.clear();
            cv_done.notify_all();
            cv_slot.notify_all();
            break;
        }
        size_t free_slot = slots.size();
        for (size_t k = 0; k < slots.size(); ++k)
            if (slots[k].inflight.empty()) {
                free_slot = k;
                break;
            }
        if (queue.empty() || free_slot == slots.size()) {
            cv_dispatch.wait(lk);
            continue;
        }
        // pack: FIFO while the outputs fit and the batch limits allow -- at most
        // max_batch_requests requests, max_batch_rays primary rays, 64 distinct grids
        // and 2^31 ray slots (64 per 8x8 tile), the limits of one batched launch
        std::vector<uint64_t> ids;
        std::vector<uint32_t> seen_grids;
        uint64_t rays = 0, ray_slots = 0;
        while (!queue.empty() && ids.size() < cfg.max_batch_requests) {
            Job &j = jobs[queue.front()];
            const uint64_t r = (uint64_t)j.req.width * j.req.height;
            if (!ids.empty() && cfg.max_batch_rays && rays + r > cfg.max_batch_rays) break;
            const uint64_t sl = (uint64_t)((j.req.width + 7u) / 8u) * ((j.req.height + 7u) / 8u) * 64u;
            if (ray_slots + sl > kMaxBatchSlots) break;  // one request alone always fits (65536^2 pixels)
            const bool new_grid =
                std::find(seen_grids.begin(), seen_grids.end(), j.req.grid) == seen_grids.end();
            if (new_grid && seen_grids.size() >= kMaxBatchGrids) break;
            uint64_t off = 0;
            if (!arena.alloc(j.bytes, &off)) break;  // wait for memory (keeps FIFO order)
            j.off = off;
            in_use += Arena::round(j.bytes);
            peak = std::max(peak, in_use);
            ids.push_back(queue.front());
            queue.pop_front();
            rays += r;
            ray_slots += sl;
            if (new_grid) seen_grids.push_back(j.req.grid);
        }
        if (ids.empty()) {  // the head request waits for memory to be released
            cv_dispatch.wait(lk);
            continue;
        }
        // grid ids of this batch -> compact launch-local indices
        std::vector<wx_grid *> bg;
        std::unordered_map<uint32_t, uint32_t> gmap;
        std::vector<wx_request> reqs(ids.size());
        const double t = now_s();
        for (size_t i = 0; i < ids.size(); ++i) {
            Job &j = jobs[ids[i]];
            j.state = Q_LAUNCHED;
            j.t_launch = t;
            j.batch_n = (uint32_t)ids.size();
            auto it = gmap.find(j.req.grid);
            if (it == gmap.end()) {
                it = gmap.emplace(j.req.grid, (uint32_t)bg.size()).first;
                bg.push_back(grids[j.req.grid]);
            }
            wx_request &r = reqs[i];
            std::memset(&r, 0, sizeof r);
            r.kind = WX_REQ_CAMERA;
            r.grid = it->second;
            r.width = j.req.width;
            r.height = j.req.height;
            r.flags = j.req.flags;
            r.pixel_offset = j.req.pixel_offset;
            r.camera = j.req.camera;
            r.light[0] = j.req.light[0], r.light[1] = j.req.light[1], r.light[2] = j.req.light[2];
            const uint64_t px = (uint64_t)j.req.width * j.req.height;
            r.out = reinterpret_cast<wx_hit *>(arena_dev + j.off);
            r.shadow_out = (j.req.flags & WX_TRACE_SHADOW) ? reinterpret_cast<wx_hit *>(arena_dev + j.off + px * 32u)
                                                           : nullptr;
        }
        Batch b;
        b.ids = ids;
        Slot &sl = slots[free_slot];
        sl.inflight.push_back(Batch());  // reserve the slot while unlocked
        st.batches++;
        batch_req_sum += (double)ids.size();
        lk.unlock();
        if (!dry) {
            b.status = wx_trace_batch(bg.data(), (uint32_t)bg.size(), reqs.data(), (uint32_t)reqs.size(), nullptr,
                                      sl.stream);
            if (hipEventCreateWithFlags(&b.ev, hipEventDisableTiming) == hipSuccess)
                (void)hipEventRecord(b.ev, sl.stream);
            else
                b.status = set_error(WX_ERR_DEVICE, "server: event create");
        }
        lk.lock();
        sl.inflight.back() = std::move(b);
        cv_slot.notify_all();
    }
}

void wx_server::complete_loop(size_t k) {
    if (!dry) (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    Slot &sl = slots[k];
    for (;;) {
        // a batch is ready once the dispatcher has filled it in (ids non-empty)
        while (!(!sl.inflight.empty() && !sl.inflight.front().ids.empty()) && !(stop && sl.inflight.empty()))
            cv_slot.wait(lk);
        if (sl.inflight.empty()) break;  // stopping and nothing in flight
        if (sl.inflight.front().ids.empty()) continue;
        Batch &b = sl.inflight.front();
        hipEvent_t ev = b.ev;
        int status = b.status;
        lk.unlock();
        if (ev) {
            if (hipEventSynchronize(ev) != hipSuccess && status == WX_OK) status = WX_ERR_DEVICE;
            (void)hipEventDestroy(ev);
        }
        const double t = now_s();
        lk.lock();
        for (uint64_t id : sl.inflight.front().ids) {
            Job &j = jobs[id];
            j.state = Q_DONE;
            j.t_done = t;
            j.status = status;
            st.completed++;
            st.primary_rays += (uint64_t)j.req.width * j.req.height;
            lat.add((t - j.t_submit) * 1e3);
        }
        sl.inflight.pop_front();
        cv_done.notify_all();
        cv_dispatch.notify_all();
        cv_slot.notify_all();
    }
}

extern "C" {

int wx_server_create(const wx_server_config *cfg, wx_server **out) {
    if (!cfg || !out) return set_error(WX_ERR_INVALID, "wx_server_create: NULL argument");
    if (cfg->streams < 1 || cfg->streams > 16) return set_error(WX_ERR_INVALID, "wx_server_create: streams in 1..16");
    if (cfg->max_batch_requests < 1 || cfg->max_batch_requests > 1024)
        return set_error(WX_ERR_INVALID, "wx_server_create: max_batch_requests in 1..1024");
    if (cfg->flags & ~(WX_SERVER_DRYRUN | WX_SERVER_RESIDENT))
        return set_error(WX_ERR_INVALID, "wx_server_create: unknown flags");
    wx_server *s = new (std::nothrow) wx_server();
    if (!s) return set_error(WX_ERR_BUDGET, "wx_server_create: host allocation failed");
    s->cfg = *cfg;
    s->dry = (cfg->flags & WX_SERVER_DRYRUN) != 0u;
    s->arena.init(cfg->hbm_budget);
    s->slots.resize(cfg->streams);
    if (!s->dry) {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
            delete s;
            return set_error(WX_ERR_DEVICE, "wx_server_create: no HIP device");
        }
        (void)hipGetDevice(&s->device);
        if (cfg->hbm_budget) {
            hipError_t e = hipMalloc(reinterpret_cast<void **>(&s->arena_dev), cfg->hbm_budget);
            if (e != hipSuccess) {
                delete s;
                return set_error(WX_ERR_BUDGET, std::string("wx_server_create: arena: ") + hipGetErrorString(e));
            }
        }
        for (Slot &sl : s->slots) {
            if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess) {
                for (Slot &o : s->slots)
                    if (o.stream) (void)hipStreamDestroy(o.stream);
                if (s->arena_dev) (void)hipFree(s->arena_dev);
                delete s;
                return set_error(WX_ERR_DEVICE, "wx_server_create: stream create");
            }
        }
    }
    s->resident = (cfg->flags & WX_SERVER_RESIDENT) != 0u && !s->dry;
    if (s->resident) {
        auto fail = [&](const char *what) {
            if (s->s_end) (void)hipEventDestroy(s->s_end);
            if (s->s_stream) (void)hipStreamDestroy(s->s_stream);
            if (s->sd) (void)hipFree(s->sd);
            if (s->mb) (void)hipHostFree(s->mb);
            for (Slot &o : s->slots)
                if (o.stream) (void)hipStreamDestroy(o.stream);
            if (s->arena_dev) (void)hipFree(s->arena_dev);
            delete s;
            return set_error(WX_ERR_DEVICE, std::string("wx_server_create: ") + what);
        };
        int cus = 0, per_cu = 0;
        if (hipHostMalloc(reinterpret_cast<void **>(&s->mb), sizeof(Mailbox), hipHostMallocCoherent) != hipSuccess) {
            s->mb = nullptr;
            return fail("pinned mailbox");
        }
        std::memset(s->mb, 0, sizeof(Mailbox));
        if (hipHostGetDevicePointer(reinterpret_cast<void **>(&s->mb_dev), s->mb, 0) != hipSuccess)
            return fail("mailbox device address");
        if (hipMalloc(reinterpret_cast<void **>(&s->sd), sizeof(SessionDev)) != hipSuccess) {
            s->sd = nullptr;
            return fail("session state");
        }
        if (hipStreamCreateWithFlags(&s->s_stream, hipStreamNonBlocking) != hipSuccess) {
            s->s_stream = nullptr;
            return fail("session stream");
        }
        if (hipEventCreateWithFlags(&s->s_end, hipEventDisableTiming) != hipSuccess) {
            s->s_end = nullptr;
            return fail("session event");
        }
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s->device) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_serve, 256, 0) != hipSuccess || cus < 1 ||
            per_cu < 1)
            return fail("session occupancy");
        s->s_blocks = cus * per_cu;  // every workgroup resident at once
    }
    if (s->resident) {
        s->dispatcher = std::thread([s] { s->resident_dispatch_loop(); });
        s->completers.emplace_back([s] { s->resident_complete_loop(); });
    } else {
        s->dispatcher = std::thread([s] { s->dispatch_loop(); });
        for (size_t k = 0; k < s->slots.size(); ++k) s->completers.emplace_back([s, k] { s->complete_loop(k); });
    }
    *out = s;
    return WX_OK;
}

int wx_server_destroy(wx_server *s) {
    if (!s) return WX_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_dispatch.notify_all();
    s->cv_slot.notify_all();
    s->dispatcher.join();
    s->cv_slot.notify_all();
    s->cv_session.notify_all();
    for (auto &t : s->completers) t.join();
    if (s->resident) {  // the completion thread ended after the last session's kernel
        (void)hipEventDestroy(s->s_end);
        (void)hipStreamDestroy(s->s_stream);
        (void)hipFree(s->sd);
        (void)hipHostFree(s->mb);
    }
    for (Slot &sl : s->slots)
        if (sl.stream) {
            wx::drop_stream_scratch(sl.stream);
            (void)hipStreamDestroy(sl.stream);
        }
    if (s->arena_dev) (void)hipFree(s->arena_dev);
    delete s;
    return WX_OK;
}

int wx_server_add_grid(wx_server *s, wx_grid *g, uint32_t *id) {
    if (!s || (!g && !s->dry) || !id) return set_error(WX_ERR_INVALID, "wx_server_add_grid: NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->resident && s->grids.size() >= kSessionGrids)
        return set_error(WX_ERR_UNSUPPORTED, "wx_server_add_grid: a resident server takes at most 64 grids");
    *id = (uint32_t)s->grids.size();
    s->grids.push_back(g);
    return WX_OK;
}

int wx_server_submit(wx_server *s, const wx_serve_request *r, uint64_t *ticket) {
    if (!s || !r || !ticket) return set_error(WX_ERR_INVALID, "wx_server_submit: NULL argument");
    if (r->flags & ~(WX_TRACE_CLIP | WX_TRACE_SHADOW)) return set_error(WX_ERR_INVALID, "wx_server_submit: bad flags");
    if (r->width == 0 || r->height == 0 || r->width > 65536u || r->height > 65536u)
        return set_error(WX_ERR_INVALID, "wx_server_submit: image size");
    std::lock_guard<std::mutex> lk(s->mu);
    if (r->grid >= s->grids.size()) return set_error(WX_ERR_INVALID, "wx_server_submit: unknown grid id");
    if (s->stop) return set_error(WX_ERR_STATE, "wx_server_submit: server is shutting down");
    if (s->cfg.max_queued && s->queue.size() >= s->cfg.max_queued) {
        s->st.rejected++;
        return set_error(WX_ERR_BUSY, "wx_server_submit: request queue full");
    }
    Job j;
    j.req = *r;
    j.bytes = job_bytes(*r);
    s->st.submitted++;
    if (Arena::round(j.bytes) > s->cfg.hbm_budget) {
        s->st.rejected++;
        return set_error(WX_ERR_BUDGET, "wx_server_submit: outputs exceed the whole HBM budget");
    }
    j.t_submit = now_s();
    const uint64_t id = s->next_id++;
    s->jobs.emplace(id, j);
    s->queue.push_back(id);
    *ticket = id;
    s->cv_dispatch.notify_all();
    return WX_OK;
}

int wx_server_wait(wx_server *s, uint64_t ticket, double timeout_s, wx_serve_result *out) {
    if (!s || !out) return set_error(WX_ERR_INVALID, "wx_server_wait: NULL argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto ready = [&] {
        auto it = s->jobs.find(ticket);
        return it == s->jobs.end() || it->second.state == Q_DONE || it->second.state == Q_CANCELLED;
    };
    if (timeout_s < 0) s->cv_done.wait(lk, ready);
    else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready))
        return set_error(WX_ERR_STATE, "wx_server_wait: timed out");
    auto it = s->jobs.find(ticket);
    if (it == s->jobs.end()) return set_error(WX_ERR_INVALID, "wx_server_wait: unknown ticket");
    const Job &j = it->second;
    if (j.state == Q_CANCELLED) return set_error(WX_ERR_STATE, "wx_server_wait: request was cancelled");
    std::memset(out, 0, sizeof *out);
    const uint64_t px = (uint64_t)j.req.width * j.req.height;
    out->primary = s->arena_dev ? reinterpret_cast<const wx_hit *>(s->arena_dev + j.off) : nullptr;
    out->shadow = (s->arena_dev && (j.req.flags & WX_TRACE_SHADOW))
                      ? reinterpret_cast<const wx_hit *>(s->arena_dev + j.off + px * 32u)
                      : nullptr;
    out->width = j.req.width;
    out->height = j.req.height;
    out->submit_s = j.t_submit;
    out->launch_s = j.t_launch;
    out->done_s = j.t_done;
    out->batch_requests = j.batch_n;
    out->status = j.status;
    return WX_OK;
}

int wx_server_release(wx_server *s, uint64_t ticket) {
    if (!s) return set_error(WX_ERR_INVALID, "wx_server_release: NULL server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(ticket);
    if (it == s->jobs.end()) return set_error(WX_ERR_INVALID, "wx_server_release: unknown ticket");
    if (it->second.state == Q_QUEUED || it->second.state == Q_LAUNCHED)
        return set_error(WX_ERR_STATE, "wx_server_release: request not done");
    if (it->second.state == Q_DONE) {
        s->arena.release(it->second.off, it->second.bytes);
        s->in_use -= Arena::round(it->second.bytes);
    }
    s->jobs.erase(it);
    s->cv_dispatch.notify_all();
    return WX_OK;
}

int wx_server_cancel(wx_server *s, uint64_t ticket) {
    if (!s) return set_error(WX_ERR_INVALID, "wx_server_cancel: NULL server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(ticket);
    if (it == s->jobs.end()) return set_error(WX_ERR_INVALID, "wx_server_cancel: unknown ticket");
    if (it->second.state != Q_QUEUED) return set_error(WX_ERR_STATE, "wx_server_cancel: already launched");
    s->queue.erase(std::find(s->queue.begin(), s->queue.end(), ticket));
    it->second.state = Q_CANCELLED;
    s->st.cancelled++;
    s->cv_done.notify_all();
    s->cv_dispatch.notify_all();
    return WX_OK;
}

int wx_server_get_stats(wx_server *s, wx_server_stats *out) {
    if (!s || !out) return set_error(WX_ERR_INVALID, "wx_server_get_stats: NULL argument");
    // copy the counters and the fixed-size histogram under the lock, read it outside
    std::unique_ptr<LatencyHist> h(new (std::nothrow) LatencyHist());
    if (!h) return set_error(WX_ERR_BUDGET, "wx_server_get_stats: host allocation failed");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        *out = s->st;
        *h = s->lat;
        out->mean_batch_requests = s->st.batches ? s->batch_req_sum / (double)s->st.batches : 0.0;
        out->budget_in_use = s->in_use;
        out->budget_peak = s->peak;
    }
    out->p50_ms = h->percentile(0.50);
    out->p99_ms = h->percentile(0.99);
    out->mean_ms = h->n ? h->sum / (double)h->n : 0.0;
    return WX_OK;
}

}  // extern "C"
