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
#include <string>
#include <vector>

#include "wx_internal.hpp"

using wx::set_error;

namespace {

constexpr uint8_t kUnknown = 255;  // the start value (the reference's MAX - 1)

// (t = 4x + 2y + z)-ordered slots of a node of side D, and where each t starts.
struct Sweep {
    std::vector<uint16_t> slots;
    std::vector<uint32_t> start;
};

Sweep make_sweep(int log) {
    const int D = 1 << log, S = D * D * D, T = 7 * (D - 1) + 1;
    Sweep w;
    w.start.assign(T + 1, 0);
    for (int s = 0; s < S; ++s) {
        const int x = s >> (2 * log), y = (s >> log) & (D - 1), z = s & (D - 1);
        w.start[4 * x + 2 * y + z + 1]++;
    }
    for (int t = 0; t < T; ++t) w.start[t + 1] += w.start[t];
    w.slots.resize(S);
    std::vector<uint32_t> fill(w.start.begin(), w.start.end() - 1);
    for (int s = 0; s < S; ++s) {
        const int x = s >> (2 * log), y = (s >> log) & (D - 1), z = s & (D - 1);
        w.slots[fill[4 * x + 2 * y + z]++] = (uint16_t)s;
    }
    return w;
}

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
                                           const int32_t *nb, const uint16_t *sweep_slots,
                                           const uint32_t *sweep_start, int forward, uint32_t *nib, uint32_t *bad) {
    constexpr int D = 1 << LOG, P = D + 2;
    constexpr uint32_t S = (uint32_t)D * D * D, W = S / 64, T = 7 * (D - 1) + 1, PS = (uint32_t)P * P * P;
    uint8_t *pad = L.pad;
    if constexpr (LOG == 3) {
        // a leaf without an inactive voxel has nothing to sweep: every slot is solid (0)
        const uint64_t *m = solid + (size_t)n * W;
        uint64_t all = ~0ull;
#pragma unroll
        for (uint32_t k = 0; k < W; ++k) all &= m[k];
        if (all == ~0ull) {
            for (uint32_t s = threadIdx.x; s < S; s += 64u) val[(size_t)n * S + s] = 0;
            if (nib) nib[(size_t)n * 64u + threadIdx.x] = 0;
            return;
        }
    }
    int32_t *s_nb = L.nb;
    if (threadIdx.x < 27u) s_nb[threadIdx.x] = nb[(size_t)n * 27 + threadIdx.x];  // the centre entry is n itself
    __syncthreads();
    // kG slots per lane at a time: their solid words and values are independent loads,
    // all in flight together (a slot's value is read even where the solid bit wins)
    constexpr uint32_t kG = 8;
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
    for (uint32_t tt = 0; tt < T; ++tt) {
        const uint32_t t = forward ? tt : T - 1u - tt;
        for (uint32_t i = sweep_start[t] + threadIdx.x; i < sweep_start[t + 1]; i += 64u) {
            const uint32_t s = sweep_slots[i];
            const int p = (int)(((s >> (2 * LOG)) + 1) * P * P + (((s >> LOG) & (D - 1)) + 1) * P + (s & (D - 1)) + 1);
            uint32_t v = pad[p];
            if (v <= 1u) continue;  // solid (0) or already 1
#pragma unroll
            for (int k = 0; k < 13; ++k) {
                const uint32_t w = pad[p + off[k]];
                if (w == 0u) {
                    v = 1u;
                    break;
                }
                const uint32_t c = w == kUnknown ? (uint32_t)kUnknown : w + 1u;
                v = c < v ? c : v;
            }
            pad[p] = (uint8_t)v;
        }
        __syncthreads();
    }
    // solid slots go back as 0; their value is never read (the solid bit answers first)
    for (uint32_t s = threadIdx.x; s < S; s += 64u)
        val[(size_t)n * S + s] = pad[((s >> (2 * LOG)) + 1) * P * P + (((s >> LOG) & (D - 1)) + 1) * P + (s & (D - 1)) + 1];
    if constexpr (LOG == 3) {
        if (nib) {  // word w = slots 8w .. 8w + 7 = (x, y) = (w >> 3, w & 7), z = 0..7
            const uint32_t w = threadIdx.x;
            const uint8_t *row = &pad[((w >> 3) + 1) * P * P + ((w & 7u) + 1) * P + 1];
            uint32_t word = 0, over = 0;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const uint32_t v = row[k];
                over |= v > 15u ? 1u : 0u;
                word |= (v & 15u) << (4 * k);
            }
            nib[(size_t)n * 64u + w] = word;
            if (over) atomicOr(bad, 1u);
        }
    }
}

// One workgroup per node of one schedule group (one launch per group).
template <int LOG>
__global__ __launch_bounds__(64) void k_sdf_nodes(const uint64_t *solid, uint8_t *val, const int32_t *nb,
                                                  const uint32_t *group, const uint16_t *sweep_slots,
                                                  const uint32_t *sweep_start, int forward, uint32_t *nib,
                                                  uint32_t *bad) {
    __shared__ NodeLds<LOG> L;
    sweep_node<LOG>(L, group[blockIdx.x], solid, val, nb, sweep_slots, sweep_start, forward, nib, bad);
}

// ---- the leaves' neighbour table and level schedules, built on the device ----
//
// Node keys go into an open-addressing hash table (keys claimed by compare-and-swap);
// the 26 neighbours of every node are looked up there. The schedule is Kahn's
// algorithm in rounds: a node's count of earlier neighbours (smaller index forward,
// larger backward) drops by one as each of them is swept, and the nodes whose count
// reaches zero form the next group -- the same groups as the longest-path levels of
// the host schedule, found while the sweep runs.

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

__global__ __launch_bounds__(64) void k_set_u32(uint32_t *p, uint32_t v) {
    if (threadIdx.x == 0) p[0] = v;
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
    if (c == 0u) front[atomicAdd(count, 1u)] = i;
}

// After the group `front` was swept: its later neighbours lose one count each; those
// reaching zero form the next group.
__global__ __launch_bounds__(256) void k_release(const int32_t *nb, const uint32_t *front, uint32_t n_front,
                                                 int forward, uint32_t *indeg, uint32_t *next, uint32_t *count) {
    const uint64_t t = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (t >= (uint64_t)n_front * 27u) return;
    const uint32_t node = front[t / 27u];
    const int32_t m = nb[(size_t)node * 27u + (uint32_t)(t % 27u)];
    if (m < 0 || (uint32_t)m == node || earlier((uint32_t)m, node, forward)) return;
    if (atomicSub(&indeg[m], 1u) == 1u) next[atomicAdd(count, 1u)] = (uint32_t)m;
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
        if (hipMemcpy(dst, p, words * 4, hipMemcpyDeviceToHost) != hipSuccess) {
            err = "download of the leaf words failed";
            return false;
        }
        return true;
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

// Kahn's algorithm in rounds on the device neighbour table: each round's group is
// swept as soon as it is known; the group sizes come back to the host to size the
// launches. (One launch per level and pass was measured slower on MI355X, both as a
// ready queue with per-node acquire/release and as Kahn rounds behind grid barriers:
// agent-scope fences write back and invalidate a whole XCD's L2.)
template <typename SweepFn>
bool sweep_rounds(uint32_t n, Ctx &c, const DevBuf &nb, SweepFn &sweep, std::string &err) {
    DevBuf indeg, fr[2], cnt;
    if (hipMalloc(&indeg.p, (size_t)n * 4) != hipSuccess || hipMalloc(&fr[0].p, (size_t)n * 4) != hipSuccess ||
        hipMalloc(&fr[1].p, (size_t)n * 4) != hipSuccess || hipMalloc(&cnt.p, 4) != hipSuccess) {
        err = "device SDF: hipMalloc of the schedule buffers failed";
        return false;
    }
    const int32_t *dnb = static_cast<const int32_t *>(nb.p);
    uint32_t *dindeg = static_cast<uint32_t *>(indeg.p), *dcnt = static_cast<uint32_t *>(cnt.p);
    const uint32_t blocks = (n + 255u) / 256u;
    auto read_count = [&](uint32_t &out) {
        return hipMemcpyAsync(&out, dcnt, 4, hipMemcpyDeviceToHost, c.stream) == hipSuccess &&
               hipStreamSynchronize(c.stream) == hipSuccess;
    };
    for (int fw = 1; fw >= 0; --fw) {  // forward pass, then backward
        uint32_t count = 0, swept = 0, cur = 0;
        hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(64), 0, c.stream, dcnt, 0u);
        hipLaunchKernelGGL(k_indegree, dim3(blocks), dim3(256), 0, c.stream, dnb, n, fw, dindeg,
                           static_cast<uint32_t *>(fr[0].p), dcnt);
        if (hipGetLastError() != hipSuccess || !read_count(count)) {
            err = "device SDF: neighbour tables failed";
            c.status = WX_ERR_DEVICE;
            return false;
        }
        while (count) {
            if (count > n - swept) {
                err = "device SDF: schedule group out of range";
                c.status = WX_ERR_DEVICE;
                return false;
            }
            const uint32_t *group = static_cast<const uint32_t *>(fr[cur].p);
            sweep(group, count, fw);
            hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(64), 0, c.stream, dcnt, 0u);
            const uint64_t rb = ((uint64_t)count * 27u + 255u) / 256u;
            hipLaunchKernelGGL(k_release, dim3((uint32_t)rb), dim3(256), 0, c.stream, dnb, group, count, fw, dindeg,
                               static_cast<uint32_t *>(fr[cur ^ 1u].p), dcnt);
            swept += count;
            cur ^= 1u;
            if (hipGetLastError() != hipSuccess || !read_count(count)) {
                err = "device SDF: sweep failed";
                c.status = WX_ERR_DEVICE;
                return false;
            }
        }
        if (swept != n) {
            err = "device SDF: the schedule did not reach every node";
            c.status = WX_ERR_DEVICE;
            return false;
        }
    }
    return true;
}

template <int LOG>
bool run_level(wx::SdfLevelJob &J, Ctx &c, std::string &err) {
    constexpr size_t S = (size_t)1 << (3 * LOG), W = S / 64;
    if (!J.nodes) return true;
    const Sweep sw = make_sweep(LOG);
    DevBuf solid, val, nb, ord0, ord1, sslots, sstart, nib, bad;
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
    if (hipMalloc(&val.p, J.nodes * S) != hipSuccess) {
        val.p = nullptr;
        err = "device SDF: hipMalloc of the distance array failed";
        return false;
    }
    if (hipMemsetAsync(val.p, kUnknown, J.nodes * S, c.stream) != hipSuccess) {
        err = "device SDF: fill failed";
        return false;
    }
    if (!upload(solid, J.solid, J.nodes * W, c.stream, err) ||
        !upload(sslots, sw.slots.data(), sw.slots.size(), c.stream, err) ||
        !upload(sstart, sw.start.data(), sw.start.size(), c.stream, err))
        return false;
    auto sweep = [&](const uint32_t *group, uint32_t n, int fw) {
        hipLaunchKernelGGL(k_sdf_nodes<LOG>, dim3(n), dim3(64), 0, c.stream, static_cast<const uint64_t *>(solid.p),
                           static_cast<uint8_t *>(val.p), static_cast<const int32_t *>(nb.p), group,
                           static_cast<const uint16_t *>(sslots.p), static_cast<const uint32_t *>(sstart.p), fw,
                           fw ? nullptr : static_cast<uint32_t *>(nib.p), static_cast<uint32_t *>(bad.p));
        ++c.launches;
    };
    if (!J.nb.empty()) {  // neighbour table and schedules from the host, one launch per group
        if (!upload(nb, J.nb.data(), J.nb.size(), c.stream, err) ||
            !upload(ord1, J.order[1].data(), J.order[1].size(), c.stream, err) ||
            !upload(ord0, J.order[0].data(), J.order[0].size(), c.stream, err))
            return false;
        for (int fw = 1; fw >= 0; --fw) {  // forward pass, then backward
            const std::vector<uint32_t> &st = J.start[fw];
            const uint32_t *ord = static_cast<const uint32_t *>(fw ? ord1.p : ord0.p);
            for (size_t g = 0; g + 1 < st.size(); ++g)
                if (st[g + 1] > st[g]) sweep(ord + st[g], st[g + 1] - st[g], fw);
        }
    } else if (!build_neighbours(J, c, nb, err) || !sweep_rounds(J.nodes, c, nb, sweep, err)) {
        return false;
    }
    uint32_t over = 0;
    if (hipGetLastError() != hipSuccess ||
        (words ? hipMemcpyAsync(&over, bad.p, 4, hipMemcpyDeviceToHost, c.stream)
               : hipMemcpyAsync(J.val, val.p, J.nodes * S, hipMemcpyDeviceToHost, c.stream)) != hipSuccess ||
        hipStreamSynchronize(c.stream) != hipSuccess) {
        err = "device SDF: kernel or download failed";
        c.status = WX_ERR_DEVICE;
        return false;
    }
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
    const bool ok = run_level<5>(jobs[0], c, err) && run_level<4>(jobs[1], c, err) && run_level<3>(jobs[2], c, err);
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
