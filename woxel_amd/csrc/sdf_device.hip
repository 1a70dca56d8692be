// Device build of the grid SDF: the two passes of VDB345::compute_sdf
// (src/vdb/vdb345.rs:290-628) on the GPU, bit-identical to the host replay.
//
// Each of the three node levels is replayed on its own. Within a pass, a node's
// slots are visited in raster order (forward) or its reverse, and a slot reads
// 13 neighbours that the pass has already visited: in the node, the slots with
// offsets (-1, dy, dz), (0, -1, dz), (0, 0, -1) (or their negatives backward); in
// the neighbouring nodes, whatever those hold at that moment. Two facts make
// this parallel without changing a value:
//  * slot (x, y, z) only reads in-node slots of smaller t = 4x + 2y + z (larger
//    t backward), so the slots of one t are independent -- a node is a sweep of
//    7 (D - 1) + 1 wavefronts (D = 8, 16, 32);
//  * the host's level schedule groups the nodes so that every neighbour a node
//    must see updated sits in an earlier group and every neighbour it must see
//    untouched in a later one -- one launch per group, one workgroup per node.
// A slot's new value is min(old, min over neighbours of (1 if the neighbour is
// solid, missing or of another level, else its value + 1)); the minimum does
// not depend on the order the neighbours are read in.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "pinned_copy.hpp"
#include "wave.hpp"
#include "wx_internal.hpp"

using wx::set_error;

namespace {

constexpr uint8_t kUnknown = 255;  // the start value (the reference's MAX - 1)

// One node swept by one workgroup (one wave). The node and a one-slot halo around
// it are staged in LDS first, coded as 0 where a read gives 1 (a solid slot, a slot
// of a missing node) and as the current value elsewhere: the neighbours a node
// reads hold still while it is swept (the schedule), so the sweep runs from LDS
// alone.
//
// Leaves (LOG 3) may also emit their packed layout words: 64 words of eight 4-bit
// codes, one word per lane, with *bad set when a code does not fit 4 bits.
template <int LOG>
struct NodeLds {
    static constexpr int D = 1 << LOG, P = D + 2;
    static constexpr uint32_t PS = (uint32_t)P * P * P;
    uint8_t pad[PS];
    int32_t nb[27];
};

template <int LOG>
__device__ __forceinline__ void sweep_node(NodeLds<LOG> &L, uint32_t n, const uint64_t *solid, uint8_t *val,
                                           const int32_t *nb, int forward, uint32_t *nib, uint32_t *bad) {
    constexpr int D = 1 << LOG, P = D + 2;
    constexpr uint32_t S = (uint32_t)D * D * D, W = S / 64, T = 7 * (D - 1) + 1, PS = (uint32_t)P * P * P;
    uint8_t *pad = L.pad;
    int32_t *s_nb = L.nb;
    // the neighbour row is loaded alongside the leaf's solid words (independent loads)
    const int32_t nb_own = threadIdx.x < 27u ? nb[(size_t)n * 27 + threadIdx.x] : -1;  // entry 13 is n itself
    if constexpr (LOG == 3) {
        // a leaf without an inactive voxel has nothing to sweep: every slot is solid (0)
        const uint64_t *m = solid + (size_t)n * W;
        uint64_t all = ~0ull;
#pragma unroll
        for (uint32_t k = 0; k < W; ++k) all &= m[k];
        if (all == ~0ull) {
            reinterpret_cast<uint64_t *>(val + (size_t)n * S)[threadIdx.x] = 0;
            if (nib) nib[(size_t)n * 64u + threadIdx.x] = 0;
            return;
        }
    }
    if (threadIdx.x < 27u) s_nb[threadIdx.x] = nb_own;
    __syncthreads();
    // kG slots per lane at a time: their solid words and values are independent loads,
    // all in flight together (a slot's value is read even where the solid bit wins); a
    // leaf's 1,000 padded slots take one round
    constexpr uint32_t kG = 16;
    for (uint32_t base = 0; base < PS; base += 64u * kG) {
        int32_t m[kG];
        uint32_t c[kG];
        uint64_t sw[kG];
        uint8_t v[kG];
#pragma unroll
        for (uint32_t k = 0; k < kG; ++k) {
            const uint32_t i = base + k * 64u + threadIdx.x;
            m[k] = -1, c[k] = 0, sw[k] = 0, v[k] = 0;
            if (i < PS) {
                int cx = (int)(i / (P * P)) - 1, cy = (int)((i / P) % P) - 1, cz = (int)(i % P) - 1;
                const int ox = cx < 0 ? -1 : (cx >= D ? 1 : 0);
                const int oy = cy < 0 ? -1 : (cy >= D ? 1 : 0);
                const int oz = cz < 0 ? -1 : (cz >= D ? 1 : 0);
                cx &= D - 1, cy &= D - 1, cz &= D - 1;
                m[k] = s_nb[(ox + 1) * 9 + (oy + 1) * 3 + (oz + 1)];
                c[k] = ((uint32_t)cx << (2 * LOG)) | ((uint32_t)cy << LOG) | (uint32_t)cz;
                if (m[k] >= 0) {
                    sw[k] = solid[(size_t)m[k] * W + (c[k] >> 6)];
                    v[k] = val[(size_t)m[k] * S + c[k]];
                }
            }
        }
#pragma unroll
        for (uint32_t k = 0; k < kG; ++k) {
            const uint32_t i = base + k * 64u + threadIdx.x;
            if (i < PS) pad[i] = (m[k] < 0 || ((sw[k] >> (c[k] & 63u)) & 1ull)) ? (uint8_t)0 : v[k];
        }
    }
    __syncthreads();
    // the 13 neighbours in padded offsets: (sgn, dy, dz), (0, sgn, dz), (0, 0, sgn), dy, dz in -1..1
    const int sgn = forward ? -1 : 1;
    int off[13];
#pragma unroll
    for (int k = 0; k < 9; ++k) off[k] = sgn * P * P + (k / 3 - 1) * P + (k % 3 - 1);
#pragma unroll
    for (int k = 0; k < 3; ++k) off[9 + k] = sgn * P + (k - 1);
    off[12] = sgn;
    // Columns (x, y) of slots (x, y, 0..D-1): lane l owns columns c = l + 64 j (x = c >> LOG,
    // y = c & (D - 1)), and on wavefront t updates slot z = t - 4x - 2y of each of them where
    // that lies in 0..D-1 (the slots of one wavefront are independent). The 13 reads are
    // independent too, and
    //   min(v, min over k of (1 if w_k == 0, 255 if w_k == 255, else w_k + 1))
    // is min(v, min_k w_k + 1) unless every w_k is 255 (then v): a solid read gives 1, and a
    // slot already at 0 or 1 keeps its value.
    constexpr uint32_t C = (uint32_t)D * D / 64u;
    for (int tt = 0; tt < (int)T; ++tt) {
        const int t = forward ? tt : (int)T - 1 - tt;
#pragma unroll
        for (uint32_t j = 0; j < C; ++j) {
            const uint32_t col = threadIdx.x + 64u * j, x = col >> LOG, y = col & (D - 1u);
            const int z = t - (int)(4u * x + 2u * y);
            if (z >= 0 && z < D) {
                const int p = (int)((x + 1u) * P * P + (y + 1u) * P + 1u) + z;
                uint32_t mw = pad[p + off[0]];
#pragma unroll
                for (int k = 1; k < 13; ++k) {
                    const uint32_t w = pad[p + off[k]];
                    mw = w < mw ? w : mw;
                }
                const uint32_t v = pad[p];
                if (mw < (uint32_t)kUnknown && mw + 1u < v) pad[p] = (uint8_t)(mw + 1u);
            }
        }
        __syncthreads();
    }
    // column c is slots c D .. c D + D - 1, stored eight at a time (solid slots go back as 0; their
    // value is never read, the solid bit answers first); a leaf lane's column is also its layout word
#pragma unroll
    for (uint32_t j = 0; j < C; ++j) {
        const uint32_t col = threadIdx.x + 64u * j, x = col >> LOG, y = col & (D - 1u);
        const int base = (int)((x + 1u) * P * P + (y + 1u) * P + 1u);
        uint64_t *dst = reinterpret_cast<uint64_t *>(val + (size_t)n * S + (size_t)col * D);
#pragma unroll
        for (int h = 0; h < D / 8; ++h) {
            uint64_t bytes = 0;
            uint32_t word = 0, over = 0;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const uint32_t v = pad[base + 8 * h + k];
                bytes |= (uint64_t)v << (8 * k);
                over |= v > 15u ? 1u : 0u;
                word |= (v & 15u) << (4 * k);
            }
            dst[h] = bytes;
            if (LOG == 3 && nib) {
                nib[(size_t)n * 64u + threadIdx.x] = word;
                if (over) atomicOr(bad, 1u);
            }
        }
    }
}

// One workgroup per node of one schedule group (one launch per group).
template <int LOG>
__global__ __launch_bounds__(64) void k_sdf_nodes(const uint64_t *solid, uint8_t *val, const int32_t *nb,
                                                  const uint32_t *group, int forward, uint32_t *nib,
                                                  uint32_t *bad) {
    __shared__ NodeLds<LOG> L;
    sweep_node<LOG>(L, group[blockIdx.x], solid, val, nb, forward, nib, bad);
}

// ---- the leaves' neighbour table and level schedules, built on the device ----
//
// Node keys go into an open-addressing hash table (keys claimed by compare-and-swap);
// the 26 neighbours of every node are looked up there. The schedule is Kahn's
// algorithm in rounds: a node's count of earlier neighbours (smaller index forward,
// larger backward) drops by one as each of them is released, and the nodes whose
// count reaches zero form the next group -- the same groups as the longest-path
// levels of the host schedule, all found before the first sweep.

constexpr uint64_t kEmptyKey = ~0ull;

__device__ __forceinline__ bool key_in_range(int32_t c) { return c >= -(1 << 20) && c < (1 << 20); }

// Three 21-bit biased fields, as the host packs node keys.
__device__ __forceinline__ uint64_t pack_key(int32_t a, int32_t b, int32_t c) {
    const int32_t bias = 1 << 20;
    return ((uint64_t)(uint32_t)(a + bias) << 42) | ((uint64_t)(uint32_t)(b + bias) << 21) |
           (uint64_t)(uint32_t)(c + bias);
}

__device__ __forceinline__ uint64_t mix_key(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    return k;
}

__global__ __launch_bounds__(256) void k_fill_u64(uint64_t *p, uint64_t n, uint64_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256u) p[i] = v;
}

// Table of 2^k >= 2 n slots: every probe sequence meets an empty slot.
__global__ __launch_bounds__(256) void k_key_put(const int32_t *keys, uint32_t n, uint64_t *kt, int32_t *vt,
                                                 uint64_t mask) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const uint64_t k = pack_key(keys[3 * (size_t)i], keys[3 * (size_t)i + 1], keys[3 * (size_t)i + 2]);
    uint64_t h = mix_key(k) & mask;
    for (;;) {
        const unsigned long long prev =
            atomicCAS(reinterpret_cast<unsigned long long *>(kt + h), (unsigned long long)kEmptyKey, k);
        if (prev == kEmptyKey || prev == k) {
            vt[h] = (int32_t)i;
            return;
        }
        h = (h + 1u) & mask;
    }
}

// nb[i * 27 + d] = index of the node at key_i + (d / 9 - 1, d / 3 % 3 - 1, d % 3 - 1), or -1.
__global__ __launch_bounds__(256) void k_neighbours(const int32_t *keys, uint32_t n, const uint64_t *kt,
                                                    const int32_t *vt, uint64_t mask, int32_t *nb) {
    const uint64_t t = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (t >= (uint64_t)n * 27u) return;
    const uint32_t i = (uint32_t)(t / 27u), d = (uint32_t)(t % 27u);
    const int32_t x = keys[3 * (size_t)i] + (int32_t)(d / 9u) - 1;
    const int32_t y = keys[3 * (size_t)i + 1] + (int32_t)(d / 3u % 3u) - 1;
    const int32_t z = keys[3 * (size_t)i + 2] + (int32_t)(d % 3u) - 1;
    int32_t r = -1;
    if (key_in_range(x) && key_in_range(y) && key_in_range(z)) {
        const uint64_t k = pack_key(x, y, z);
        for (uint64_t h = mix_key(k) & mask; kt[h] != kEmptyKey; h = (h + 1u) & mask)
            if (kt[h] == k) {
                r = vt[h];
                break;
            }
    }
    nb[t] = r;
}

__device__ __forceinline__ bool earlier(uint32_t m, uint32_t n, int forward) { return forward ? m < n : m > n; }

// indeg[i] = neighbours swept before i in this pass; the nodes without any form group 0.
__global__ __launch_bounds__(256) void k_indegree(const int32_t *nb, uint32_t n, int forward, uint32_t *indeg,
                                                  uint32_t *front, uint32_t *count) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    uint32_t c = 0;
    for (uint32_t d = 0; d < 27u; ++d) {
        const int32_t m = nb[(size_t)i * 27u + d];
        c += (m >= 0 && (uint32_t)m != i && earlier((uint32_t)m, i, forward)) ? 1u : 0u;
    }
    indeg[i] = c;
    if (c == 0u) {
        const uint32_t k = atomicAdd(count, 1u);
        if (k < n) front[k] = i;
    }
}

// One Kahn round of both passes' schedules, computed ahead of the sweeps (a node's
// group depends only on the neighbour table). Pass p keeps its groups one after the
// other in order[p] and group r's {start, size} in rs[p][r]: the round releases group
// r -- its later neighbours lose one count each, those reaching zero are appended as
// group r + 1 -- and records where group r + 1 starts. Blocks [0, G) take the forward
// pass, [G, 2G) the backward one; each strides over its group's nodes, one node per
// lane, so the launch shape does not depend on the group size and no size has to come
// back to the host between rounds. Appends past n (impossible for a neighbour table of
// this file's making) are dropped rather than written out of bounds.
struct Schedule {
    const int32_t *nb;
    uint32_t n;
    uint32_t *indeg[2], *order[2];
    uint2 *rs[2];
};

// A lane takes one node of the group: its 27 neighbour entries are loaded together and
// its later neighbours' counts are decremented with all the atomics in flight at once
// (the results are compared only after the last one is issued), so a round costs one
// chain of dependent accesses -- node, row, counts, append -- per lane. The lanes'
// releases are appended with one counter update per wave (a group's order does not
// change a value).
__global__ __launch_bounds__(256) void k_release_round(Schedule s, uint32_t r) {
    const uint32_t G = gridDim.x / 2u, p = blockIdx.x < G ? 0u : 1u;
    const int forward = p == 0u ? 1 : 0;
    const uint2 g = s.rs[p][r];  // group r: start, size
    const bool ok = g.x <= s.n && g.y <= s.n - g.x;  // else the host finds the records inconsistent
    const uint32_t end = ok ? g.x + g.y : s.n, size = ok ? g.y : 0u, room = s.n - end;
    if (blockIdx.x % G == 0u && threadIdx.x == 0u) s.rs[p][r + 1u].x = end;
    uint32_t *cnt = &s.rs[p][r + 1u].y, *indeg = s.indeg[p], *order = s.order[p];
    // j0 is the same for every lane of the block, so every wave runs each trip whole (the scan
    // below needs all 64 lanes)
    for (uint64_t j0 = (uint64_t)(blockIdx.x % G) * 256u; j0 < size; j0 += (uint64_t)G * 256u) {
        const uint64_t j = j0 + threadIdx.x;
        int32_t m[27];
        uint32_t old[27];
        if (j < size) {
            const uint32_t node = order[g.x + (uint32_t)j];
            const int32_t *row = s.nb + (size_t)node * 27u;
#pragma unroll
            for (int d = 0; d < 27; ++d) m[d] = row[d];
            // the whole row is in before the first atomic is issued (the compiler would
            // otherwise wait for a late row load behind the atomics issued so far)
            asm volatile("" : "+v"(m[0]), "+v"(m[1]), "+v"(m[2]), "+v"(m[3]), "+v"(m[4]), "+v"(m[5]), "+v"(m[6]),
                         "+v"(m[7]), "+v"(m[8]), "+v"(m[9]), "+v"(m[10]), "+v"(m[11]), "+v"(m[12]), "+v"(m[13]));
            asm volatile("" : "+v"(m[14]), "+v"(m[15]), "+v"(m[16]), "+v"(m[17]), "+v"(m[18]), "+v"(m[19]),
                         "+v"(m[20]), "+v"(m[21]), "+v"(m[22]), "+v"(m[23]), "+v"(m[24]), "+v"(m[25]), "+v"(m[26]));
#pragma unroll
            for (int d = 0; d < 27; ++d) {
                // later in this pass's sweep order (a missing neighbour, -1, is never later)
                const uint32_t u = (uint32_t)m[d];
                old[d] = 0u;
                if (m[d] >= 0 && (forward ? u > node : u < node)) old[d] = atomicSub(&indeg[u], 1u);
            }
        } else {
#pragma unroll
            for (int d = 0; d < 27; ++d) m[d] = -1, old[d] = 0u;
        }
        uint32_t rel = 0;  // bit d: neighbour d's count reached zero
#pragma unroll
        for (int d = 0; d < 27; ++d) rel |= (old[d] == 1u ? 1u : 0u) << d;
        uint32_t total;
        const uint32_t off = wxw::wave_scan((uint32_t)__builtin_popcount(rel), total);
        uint32_t base = 0;
        if (__lane_id() == 0u && total) base = atomicAdd(cnt, total);
        uint32_t k = __shfl(base, 0) + off;
#pragma unroll
        for (int d = 0; d < 27; ++d)
            if ((rel >> d) & 1u) {
                if (k < room) order[end + k] = (uint32_t)m[d];
                ++k;
            }
    }
}

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

// Leaf words left in device memory by the replay, freed with the tree.
struct HipWords final : wx::DeviceWords {
    void *p;
    int dev;
    HipWords(void *ptr, int device) : p(ptr), dev(device) {}
    ~HipWords() override {
        int cur = 0;
        if (hipGetDevice(&cur) != hipSuccess) return;
        if (cur != dev) (void)hipSetDevice(dev);
        (void)hipFree(p);
        if (cur != dev) (void)hipSetDevice(cur);
    }
    const void *data() const override { return p; }
    int device() const override { return dev; }
    bool download(uint32_t *dst, size_t words, std::string &err) const override {
        int cur = 0;
        if (hipGetDevice(&cur) != hipSuccess || (cur != dev && hipSetDevice(dev) != hipSuccess)) {
            err = "download of the leaf words: device";
            return false;
        }
        bool ok;
        if (words * 4 >= wx::PinnedCopier::kMin) {  // the copier's stream on the words' device
            wx::PinnedCopier c;
            ok = c.ok() && c.to_host(dst, p, words * 4);
        } else {
            ok = hipMemcpy(dst, p, words * 4, hipMemcpyDeviceToHost) == hipSuccess;
        }
        if (cur != dev) (void)hipSetDevice(cur);
        if (!ok) err = "download of the leaf words failed";
        return ok;
    }
};

struct Ctx {
    hipStream_t stream;
    int status = WX_OK;
    uint64_t launches = 0;
    double t_enter = 0, t_device = 0;  // when the replay was handed the prepared levels; time on the device
};

double seconds() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Phase timings of one level's device replay on stderr when WX_TRACE_BUILD is set
// (the host build's phase lines, one level deeper).
struct LevelTrace {
    bool on = std::getenv("WX_TRACE_BUILD") != nullptr;
    int log;
    double last = seconds();
    explicit LevelTrace(int l) : log(l) {}
    void phase(const char *name) {
        if (!on) return;
        const double t = seconds();
        std::fprintf(stderr, "wx build:   level %d %-14s %8.3f s\n", log, name, t - last);
        last = t;
    }
};

template <typename T>
bool upload(DevBuf &b, const T *src, size_t n, hipStream_t st, std::string &err) {
    const size_t bytes = (n ? n : 1) * sizeof(T);
    if (hipMalloc(&b.p, bytes) != hipSuccess) {
        b.p = nullptr;
        err = "device SDF: hipMalloc of " + std::to_string(bytes) + " bytes failed";
        return false;
    }
    if (n && hipMemcpyAsync(b.p, src, n * sizeof(T), hipMemcpyHostToDevice, st) != hipSuccess) {
        err = "device SDF: upload failed";
        return false;
    }
    return true;
}

// The leaves' neighbour table on the device (from J.keys, DFS order) into nb.
bool build_neighbours(wx::SdfLevelJob &J, Ctx &c, DevBuf &nb, std::string &err) {
    const uint32_t n = J.nodes;
    uint64_t cap = 64;
    while (cap < 2ull * n) cap <<= 1;
    DevBuf keys, kt, vt;
    if (!J.keys || !upload(keys, J.keys, (size_t)n * 3, c.stream, err)) {
        if (!J.keys) err = "device SDF: no neighbour table and no node keys";
        return false;
    }
    if (hipMalloc(&kt.p, cap * 8) != hipSuccess || hipMalloc(&vt.p, cap * 4) != hipSuccess ||
        hipMalloc(&nb.p, (size_t)n * 27 * 4) != hipSuccess) {
        err = "device SDF: hipMalloc of the neighbour tables failed";
        return false;
    }
    const uint32_t blocks = (n + 255u) / 256u;
    const uint64_t nb_blocks = ((uint64_t)n * 27u + 255u) / 256u;
    hipLaunchKernelGGL(k_fill_u64, dim3(4096), dim3(256), 0, c.stream, static_cast<uint64_t *>(kt.p), cap, kEmptyKey);
    hipLaunchKernelGGL(k_key_put, dim3(blocks), dim3(256), 0, c.stream, static_cast<const int32_t *>(keys.p), n,
                       static_cast<uint64_t *>(kt.p), static_cast<int32_t *>(vt.p), cap - 1u);
    hipLaunchKernelGGL(k_neighbours, dim3((uint32_t)nb_blocks), dim3(256), 0, c.stream,
                       static_cast<const int32_t *>(keys.p), n, static_cast<const uint64_t *>(kt.p),
                       static_cast<const int32_t *>(vt.p), cap - 1u, static_cast<int32_t *>(nb.p));
    // the tables (keys, kt, vt) are freed when this returns: wait for their readers
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c.stream) != hipSuccess) {
        err = "device SDF: neighbour table failed";
        c.status = WX_ERR_DEVICE;
        return false;
    }
    return true;
}

// Kahn's algorithm in rounds on the device neighbour table, for both passes at once
// and ahead of the sweeps, on a worker thread with a stream of its own: the rounds
// are launched in batches (64, 128, ... up to 1,024 rounds) and each batch's group
// records come back when it ends -- rounds past the end are empty launches -- so the
// caller sweeps the forward pass's groups batch by batch while later rounds run, and
// uploads the solid masks meanwhile. (One launch per level and pass was measured
// slower on MI355X, both as a ready queue with per-node acquire/release and as Kahn
// rounds behind grid barriers: agent-scope fences write back and invalidate a whole
// XCD's L2.)
class Scheduler {
  public:
    ~Scheduler() {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_ = true;
        }
        if (th_.joinable()) th_.join();
        if (stream_) (void)hipStreamDestroy(stream_);
    }

    // Allocates, computes the in-degrees and starts the rounds (nb must be complete).
    bool start(uint32_t n, const DevBuf &nb, std::string &err) {
        n_ = n;
        if (hipGetDevice(&dev_) != hipSuccess || hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking) != hipSuccess) {
            err = "device SDF: schedule stream failed";
            return false;
        }
        for (int p = 0; p < 2; ++p)
            if (hipMalloc(&indeg_[p].p, (size_t)n * 4) != hipSuccess || hipMalloc(&order_[p].p, (size_t)n * 4) != hipSuccess) {
                err = "device SDF: hipMalloc of the schedule buffers failed";
                return false;
            }
        if (!grow(4096, 0, err)) return false;
        s_.nb = static_cast<const int32_t *>(nb.p);
        s_.n = n;
        for (int p = 0; p < 2; ++p) {
            s_.indeg[p] = static_cast<uint32_t *>(indeg_[p].p);
            s_.order[p] = static_cast<uint32_t *>(order_[p].p);
            hipLaunchKernelGGL(k_indegree, dim3((n + 255u) / 256u), dim3(256), 0, stream_, s_.nb, n, p == 0 ? 1 : 0,
                               s_.indeg[p], s_.order[p], &s_.rs[p][0].y);
        }
        if (hipGetLastError() != hipSuccess) {
            err = "device SDF: in-degree launch failed";
            return false;
        }
        th_ = std::thread([this] { run(); });
        return true;
    }

    // Waits for records past `have` of pass p (0 forward, 1 backward) and appends them
    // to out; false once every record is out (or on failure: see failed()).
    bool next(int p, std::vector<uint2> &out) {
        std::unique_lock<std::mutex> g(mu_);
        cv_.wait(g, [&] { return known_ > out.size() || done_ || failed_; });
        if (failed_) return false;
        out.insert(out.end(), rec_[p].begin() + (ptrdiff_t)out.size(), rec_[p].begin() + (ptrdiff_t)known_);
        return !(done_ && out.size() == known_);
    }

    bool failed(std::string &err) {
        if (th_.joinable()) th_.join();
        if (failed_) err = err_;
        return failed_;
    }
    const uint32_t *order(int p) const { return s_.order[p]; }

  private:
    bool grow(uint64_t want, uint64_t used, std::string &err) {
        uint64_t nc = cap_ ? cap_ : 4096;
        while (nc < want) nc *= 2;
        for (int p = 0; p < 2; ++p) {
            DevBuf nr;
            if (hipMalloc(&nr.p, (nc + 1) * sizeof(uint2)) != hipSuccess ||
                hipMemsetAsync(nr.p, 0, (nc + 1) * sizeof(uint2), stream_) != hipSuccess ||
                (used && hipMemcpyAsync(nr.p, rs_[p].p, used * sizeof(uint2), hipMemcpyDeviceToDevice, stream_) !=
                             hipSuccess) ||
                hipStreamSynchronize(stream_) != hipSuccess) {  // the old records are freed below
                err = "device SDF: schedule records: allocation or copy failed";
                return false;
            }
            std::swap(rs_[p].p, nr.p);
            s_.rs[p] = static_cast<uint2 *>(rs_[p].p);
        }
        cap_ = nc;
        return true;
    }

    void fail(const std::string &e) {
        std::lock_guard<std::mutex> g(mu_);
        failed_ = true;
        err_ = e;
        cv_.notify_all();
    }

    void run() {
        if (hipSetDevice(dev_) != hipSuccess) return fail("device SDF: schedule thread: device");
        // 128 workgroups per pass: measured on the 28 M-leaf isosurface's schedule (4,032 rounds,
        // run beside the forward sweeps), the device part took 0.324 s against 0.328-0.331 s
        // at 64 and 256, 0.343 s at 512 and 0.394 s at 2,048 (larger grids cost more to
        // dispatch than the strided loop over a group's (node, neighbour) pairs saves)
        constexpr uint32_t kBlocksPerPass = 128;
        uint64_t r = 0;  // rounds launched; records 0 .. r are final after each batch
        std::vector<uint2> got[2];
        for (uint64_t batch = 64;; batch = batch < 1024 ? batch * 2 : batch) {
            {
                std::lock_guard<std::mutex> g(mu_);
                if (stop_) return;
            }
            std::string e;
            if (r + batch > cap_ && !grow(r + batch, r + 1, e)) return fail(e);
            const uint64_t r0 = r;
            for (uint64_t k = 0; k < batch; ++k, ++r)
                hipLaunchKernelGGL(k_release_round, dim3(2 * kBlocksPerPass), dim3(256), 0, stream_, s_, (uint32_t)r);
            for (int p = 0; p < 2; ++p) {
                got[p].resize(r + 1);
                if (hipMemcpyAsync(got[p].data() + r0, &s_.rs[p][r0], (r - r0 + 1) * sizeof(uint2),
                                   hipMemcpyDeviceToHost, stream_) != hipSuccess)
                    return fail("device SDF: schedule download failed");
            }
            if (hipGetLastError() != hipSuccess || hipStreamSynchronize(stream_) != hipSuccess)
                return fail("device SDF: schedule rounds failed");
            // group r is empty in both passes: so is every later one
            const bool end = got[0][r].y == 0 && got[1][r].y == 0;
            if (!end && r > (uint64_t)n_) return fail("device SDF: the schedule did not drain");
            std::lock_guard<std::mutex> g(mu_);
            for (int p = 0; p < 2; ++p) rec_[p].assign(got[p].begin(), got[p].end());
            known_ = r + 1;
            done_ = end;
            cv_.notify_all();
            if (end) return;
        }
    }

    uint32_t n_ = 0;
    int dev_ = 0;
    hipStream_t stream_ = nullptr;
    DevBuf indeg_[2], order_[2], rs_[2];
    uint64_t cap_ = 0;
    Schedule s_{};
    std::thread th_;
    std::mutex mu_;
    std::condition_variable cv_;
    std::vector<uint2> rec_[2];
    uint64_t known_ = 0;
    bool done_ = false, failed_ = false, stop_ = false;
    std::string err_;
};

// Sweeps pass p's groups as their records arrive: consecutive groups, each inside the
// node range (checked before its launch), covering every node exactly once.
template <typename SweepFn>
bool sweep_pass(Scheduler &sch, int p, uint32_t n, SweepFn &sweep, Ctx &c, std::string &err) {
    std::vector<uint2> rec;
    size_t used = 0;
    uint64_t at = 0;
    bool more = true;
    while (more) {
        more = sch.next(p, rec);
        for (; used < rec.size(); ++used) {
            const uint2 g = rec[used];
            if (g.x != at || g.y > n - g.x) {
                err = "device SDF: schedule groups out of order";
                c.status = WX_ERR_DEVICE;
                return false;
            }
            at += g.y;
            if (g.y) sweep(sch.order(p) + g.x, g.y, p == 0 ? 1 : 0);
        }
        if (!more && sch.failed(err)) {
            c.status = WX_ERR_DEVICE;
            return false;
        }
    }
    if (at != n) {
        err = "device SDF: the schedule did not reach every node";
        c.status = WX_ERR_DEVICE;
        return false;
    }
    return true;
}

template <int LOG>
bool run_level(wx::SdfLevelJob &J, Ctx &c, std::string &err) {
    constexpr size_t S = (size_t)1 << (3 * LOG), W = S / 64;
    if (!J.nodes) return true;
    LevelTrace tr(LOG);
    DevBuf solid, val, nb, ord0, ord1, nib, bad;
    const bool words = LOG == 3 && J.words;  // leaves leave as packed layout words, kept on the device
    if (words) {
        const uint32_t zero = 0;
        if (hipMalloc(&nib.p, (size_t)J.nodes * 256) != hipSuccess || hipMalloc(&bad.p, 4) != hipSuccess) {
            err = "device SDF: hipMalloc of the leaf words failed";
            return false;
        }
        if (hipMemcpyAsync(bad.p, &zero, 4, hipMemcpyHostToDevice, c.stream) != hipSuccess ||
            hipStreamSynchronize(c.stream) != hipSuccess) {
            err = "device SDF: flag reset failed";
            return false;
        }
    }
    // every slot starts unknown (255): filled on the device instead of copied
    auto make_val = [&] {
        if (hipMalloc(&val.p, J.nodes * S) != hipSuccess) {
            val.p = nullptr;
            err = "device SDF: hipMalloc of the distance array failed";
            return false;
        }
        if (hipMemsetAsync(val.p, kUnknown, J.nodes * S, c.stream) != hipSuccess) {
            err = "device SDF: fill failed";
            return false;
        }
        return true;
    };
    // the solid masks come from pageable host memory; a large set goes through pinned
    // staging on a stream of the copier's own (host threads fill one buffer while the DMA
    // drains the other) and is on the device when this returns
    const size_t solid_bytes = J.nodes * W * sizeof(uint64_t);
    const bool staged = solid_bytes >= wx::PinnedCopier::kMin;
    auto staged_masks = [&](std::string &e) {
        if (hipMalloc(&solid.p, solid_bytes) != hipSuccess) {
            solid.p = nullptr;
            e = "device SDF: hipMalloc of the solid masks failed";
            return false;
        }
        wx::PinnedCopier cp;
        if (!cp.ok() || !cp.to_device(solid.p, J.solid, solid_bytes)) {
            e = "device SDF: upload of the solid masks failed";
            return false;
        }
        return true;
    };
    auto prepare = [&] {
        if (!make_val()) return false;
        return staged ? staged_masks(err) : upload(solid, J.solid, J.nodes * W, c.stream, err);
    };
    std::unique_ptr<Scheduler> sch;  // outlives the sweeps, which read its groups
    const uint64_t *sp = nullptr;     // the solid masks the sweeps read: solid, or the tree's device copy
    auto sweep = [&](const uint32_t *group, uint32_t n, int fw) {
        hipLaunchKernelGGL(k_sdf_nodes<LOG>, dim3(n), dim3(64), 0, c.stream, sp,
                           static_cast<uint8_t *>(val.p), static_cast<const int32_t *>(nb.p), group,
                           fw,
                           fw ? nullptr : static_cast<uint32_t *>(nib.p), static_cast<uint32_t *>(bad.p));
        ++c.launches;
    };
    if (!J.nb.empty()) {  // neighbour table and schedules from the host, one launch per group
        if (!prepare() || !upload(nb, J.nb.data(), J.nb.size(), c.stream, err) ||
            !upload(ord1, J.order[1].data(), J.order[1].size(), c.stream, err) ||
            !upload(ord0, J.order[0].data(), J.order[0].size(), c.stream, err))
            return false;
        sp = static_cast<const uint64_t *>(solid.p);
        tr.phase("uploads");
        for (int fw = 1; fw >= 0; --fw) {  // forward pass, then backward
            const std::vector<uint32_t> &st = J.start[fw];
            const uint32_t *ord = static_cast<const uint32_t *>(fw ? ord1.p : ord0.p);
            for (size_t g = 0; g + 1 < st.size(); ++g)
                if (st[g + 1] > st[g]) sweep(ord + st[g], st[g + 1] - st[g], fw);
        }
    } else {  // the schedule on the device, its rounds overlapping the upload and the forward sweeps
        sch.reset(new Scheduler);
        // a large mask set goes up on a thread of its own while the neighbour table is built
        // and the rounds start (it has its own stream; the sweeps are launched after the join)
        int dev = 0;
        (void)hipGetDevice(&dev);
        // masks a device generator left on this device are read where they are
        const bool resident = J.solid_dev && J.solid_dev_device == dev;
        std::string up_err;
        bool up_ok = true;
        std::thread up;
        if (staged && !resident)
            up = std::thread([&] {
                try {  // nothing may escape a thread: an exception becomes this call's error
                    up_ok = hipSetDevice(dev) == hipSuccess ? staged_masks(up_err)
                                                            : (up_err = "device SDF: upload thread: device", false);
                } catch (...) {
                    up_ok = false;
                    up_err = "device SDF: upload of the solid masks failed (host resources)";
                }
            });
        bool ok = build_neighbours(J, c, nb, err);
        if (ok) {
            tr.phase("neighbours");
            ok = sch->start(J.nodes, nb, err);
        }
        if (up.joinable()) up.join();
        if (!ok) return false;
        if (!up_ok) {
            err = up_err;
            return false;
        }
        tr.phase("rounds started");
        if (!(staged || resident ? make_val() : prepare())) return false;
        sp = resident ? static_cast<const uint64_t *>(J.solid_dev) : static_cast<const uint64_t *>(solid.p);
        tr.phase(resident ? "masks resident" : "masks up");
        if (!sweep_pass(*sch, 0, (uint32_t)J.nodes, sweep, c, err)) return false;
        tr.phase("forward queued");
        if (!sweep_pass(*sch, 1, (uint32_t)J.nodes, sweep, c, err)) return false;
        tr.phase("backward queued");
    }
    uint32_t over = 0;
    tr.phase("launched");
    if (hipGetLastError() != hipSuccess ||
        (words ? hipMemcpyAsync(&over, bad.p, 4, hipMemcpyDeviceToHost, c.stream)
               : hipMemcpyAsync(J.val, val.p, J.nodes * S, hipMemcpyDeviceToHost, c.stream)) != hipSuccess ||
        hipStreamSynchronize(c.stream) != hipSuccess) {
        err = "device SDF: kernel or download failed";
        c.status = WX_ERR_DEVICE;
        return false;
    }
    tr.phase("done");
    if (over) {
        err = "compute_sdf: voxel distance above 15 (unsupported topology)";
        return false;
    }
    if (words) {  // the words stay where the device grid will be built
        int dev = 0;
        (void)hipGetDevice(&dev);
        J.kept.reset(new HipWords(nib.p, dev));
        nib.p = nullptr;
    }
    return true;
}

bool run_device(wx::SdfLevelJob *jobs, void *ctx, std::string &err) {
    Ctx &c = *static_cast<Ctx *>(ctx);
    c.t_enter = seconds();
    LevelTrace tr(0);
    const bool ok = run_level<5>(jobs[0], c, err) && run_level<4>(jobs[1], c, err) && run_level<3>(jobs[2], c, err);
    tr.phase("all, freed");
    c.t_device = seconds() - c.t_enter;
    return ok;
}

}  // namespace

extern "C" int wx_vdb_finalize_device(wx_vdb *v, int threads, wx_stream stream, wx_build_info *info) {
    if (!v) return set_error(WX_ERR_INVALID, "wx_vdb_finalize_device: NULL tree");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        return set_error(WX_ERR_DEVICE, "wx_vdb_finalize_device: no HIP device");
    try {
        Ctx c;
        c.stream = static_cast<hipStream_t>(stream);
        std::string err;
        const double t0 = seconds();
        if (!v->tree.finalize_with(threads < 1 ? 1 : threads, run_device, &c, err))
            return set_error(c.status != WX_OK ? c.status : WX_ERR_UNSUPPORTED, "wx_vdb_finalize_device: " + err);
        if (info) {
            info->launches = c.launches;
            info->total_s = seconds() - t0;
            info->sdf_device_s = c.t_device;
            info->before_sdf_s = c.t_enter > 0 ? c.t_enter - t0 : 0.0;
        }
        return WX_OK;
    } catch (const std::bad_alloc &) {
        return set_error(WX_ERR_BUDGET, "wx_vdb_finalize_device: host allocation failed");
    }
}
