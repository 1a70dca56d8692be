// Device generator of the value-noise isosurface leaves: the voxels of
// HostVdb::add_noise, computed by the GPU in slabs of candidate leaves and appended
// to the host tree. The noise, the leaf classes and the threshold are the shared
// definitions of noise.hpp, so both generators produce the same leaves bit for bit.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#define WX_NOISE_FN __host__ __device__ inline
#include "noise.hpp"
#include "pinned_copy.hpp"
#include "wave.hpp"
#include "wx_internal.hpp"

using wx::set_error;

namespace {

constexpr uint64_t kSlab = 1ull << 24;  // candidate leaves per launch (1.3 GB of output at most)

struct LeafBox {
    int32_t l0[3];  // first leaf key
    uint32_t ny, nz;
};

// Candidate c = first + i of the x-major leaf box: its leaf key.
__device__ __forceinline__ void candidate_key(const LeafBox &bx, uint64_t c, int32_t &kx, int32_t &ky, int32_t &kz) {
    const uint64_t r = c / bx.nz;
    kx = bx.l0[0] + (int32_t)(r / bx.ny);
    ky = bx.l0[1] + (int32_t)(r % bx.ny);
    kz = bx.l0[2] + (int32_t)(c % bx.nz);
}

using wxw::wave_scan;

// Leaf masks kept in device memory for the device SDF replay; the tree frees them.
struct HipMasks final : wx::DeviceBuffer {
    void *p;
    int dev;
    HipMasks(void *ptr, int device) : p(ptr), dev(device) {}
    ~HipMasks() override {
        int cur = 0;
        if (hipGetDevice(&cur) != hipSuccess) return;
        if (cur != dev) (void)hipSetDevice(dev);
        (void)hipFree(p);
        if (cur != dev) (void)hipSetDevice(cur);
    }
    const void *data() const override { return p; }
    int device() const override { return dev; }
};

constexpr uint32_t kPerLane = 8;  // candidates per lane (classify), listed leaves per wave (voxels)

// Lane t takes candidates i = 8t .. 8t + 7 (c = first + i, consecutive along z of the
// x-major leaf box) and classes them by their centre value. A leaf with every voxel
// active takes an output slot (mask all ones); one to decide voxel by voxel goes to
// the work list `todo`; an empty one ends here. One counter update per list and wave.
__global__ __launch_bounds__(256) void k_noise_classify(wxnoise::Params p, LeafBox bx, uint64_t first, uint32_t count,
                                                        uint32_t *counter, int32_t *keys, uint8_t *masks,
                                                        uint32_t *todo, uint32_t *ntodo) {
    const uint64_t i0 = ((uint64_t)blockIdx.x * 256u + threadIdx.x) * kPerLane;
    if (((uint64_t)blockIdx.x * 256u + (threadIdx.x & ~63u)) * kPerLane >= count) return;  // the whole wave
    int32_t kx, ky, kz;
    candidate_key(bx, first + i0, kx, ky, kz);
    uint32_t cls = 0, n1 = 0, n2 = 0;  // 2 bits per candidate
    {
        int32_t x = kx, y = ky, z = kz;
        for (uint32_t k = 0; k < kPerLane; ++k) {
            if (i0 + k < count) {
                const int32_t b[3] = {x * 8, y * 8, z * 8};
                bool inside;
                const uint32_t c = (uint32_t)wxnoise::leaf_class(p, b, inside);
                cls |= c << (2u * k);
                n1 += c == 1u ? 1u : 0u;
                n2 += c == 2u ? 1u : 0u;
            }
            if (++z == bx.l0[2] + (int32_t)bx.nz) {  // the next candidate of the x-major box
                z = bx.l0[2];
                if (++y == bx.l0[1] + (int32_t)bx.ny) y = bx.l0[1], ++x;
            }
        }
    }
    uint32_t tot = 0;
    const uint32_t pre = wave_scan(n1 | n2 << 16, tot);
    uint32_t base1 = 0, base2 = 0;
    if (__lane_id() == 0) {
        if (tot & 0xFFFFu) base1 = atomicAdd(counter, tot & 0xFFFFu);
        if (tot >> 16) base2 = atomicAdd(ntodo, tot >> 16);
    }
    base1 = __shfl(base1, 0) + (pre & 0xFFFFu);
    base2 = __shfl(base2, 0) + (pre >> 16);
    int32_t x = kx, y = ky, z = kz;
    for (uint32_t k = 0; k < kPerLane; ++k) {
        const uint32_t c = (cls >> (2u * k)) & 3u;
        if (c == 1u) {
            uint4 *m = reinterpret_cast<uint4 *>(masks + (size_t)base1 * 64u);
            const uint4 ones = make_uint4(~0u, ~0u, ~0u, ~0u);
            m[0] = ones, m[1] = ones, m[2] = ones, m[3] = ones;
            keys[(size_t)base1 * 3u] = x, keys[(size_t)base1 * 3u + 1u] = y, keys[(size_t)base1 * 3u + 2u] = z;
            ++base1;
        } else if (c == 2u) {
            todo[base2++] = (uint32_t)(i0 + k);
        }
        if (++z == bx.l0[2] + (int32_t)bx.nz) {
            z = bx.l0[2];
            if (++y == bx.l0[1] + (int32_t)bx.ny) y = bx.l0[1], ++x;
        }
    }
}

// One wave per 8 listed candidates, decided one after the other: lane (x, y) = (lane >> 3,
// lane & 7) decides the voxels (x, y, 0..7), which are bits 0..7 of byte y of mask word x:
// byte `lane` of the leaf's 64-byte mask. The leaves with an active voxel take output
// slots, one counter update per wave.
__global__ __launch_bounds__(256) void k_noise_voxels(wxnoise::Params p, LeafBox bx, uint64_t first,
                                                      const uint32_t *todo, uint32_t ntodo, uint32_t *counter,
                                                      int32_t *keys, uint8_t *masks) {
    const uint32_t w0 = (blockIdx.x * 4u + (threadIdx.x >> 6)) * kPerLane;
    if (w0 >= ntodo) return;  // the whole wave
    const uint32_t lane = threadIdx.x & 63u;
    // the lane's column lies in one lattice cell per octave but where it crosses a cell
    // border: each cell's corners are hashed once
    wxnoise::Cell cache[wxnoise::kOctaves];
    wxnoise::clear_cache(cache);
    uint32_t bytes[kPerLane];
    uint32_t keep = 0;  // wave-uniform: bit k set when listed leaf k has an active voxel
#pragma unroll
    for (uint32_t k = 0; k < kPerLane; ++k) {
        bytes[k] = 0;
        if (w0 + k < ntodo) {
            int32_t kx, ky, kz;
            candidate_key(bx, first + todo[w0 + k], kx, ky, kz);
            const int32_t x = kx * 8 + (int32_t)(lane >> 3), y = ky * 8 + (int32_t)(lane & 7u);
            for (int z = 0; z < 8; ++z)
                if (wxnoise::voxel_active(p, x, y, kz * 8 + z, cache)) bytes[k] |= 1u << z;
            if (__ballot(bytes[k] != 0u) != 0) keep |= 1u << k;
        }
    }
    if (!keep) return;
    uint32_t slot = 0;
    if (lane == 0) slot = atomicAdd(counter, (uint32_t)__builtin_popcount(keep));
    slot = __shfl(slot, 0);
#pragma unroll
    for (uint32_t k = 0; k < kPerLane; ++k) {
        if (!((keep >> k) & 1u)) continue;
        masks[(size_t)slot * 64u + lane] = (uint8_t)bytes[k];
        if (lane < 3u) {
            int32_t kx, ky, kz;
            candidate_key(bx, first + todo[w0 + k], kx, ky, kz);
            keys[(size_t)slot * 3u + lane] = lane == 0u ? kx : (lane == 1u ? ky : kz);
        }
        ++slot;
    }
}

// The DFS position of a leaf (the key the tree sorts its leaves by): the node-5 key
// fields biased to 12 bits each, the node-4 offset in its node-5, the leaf offset in
// its node-4.
__device__ __forceinline__ uint64_t dfs_key(int32_t x, int32_t y, int32_t z) {
    const uint64_t off4 = ((uint64_t)((x >> 4) & 31) << 10) | ((uint64_t)((y >> 4) & 31) << 5) | (uint64_t)((z >> 4) & 31);
    const uint64_t off3 = ((uint64_t)(x & 15) << 8) | ((uint64_t)(y & 15) << 4) | (uint64_t)(z & 15);
    return ((uint64_t)((x >> 9) + 2048) << 51) | ((uint64_t)((y >> 9) + 2048) << 39) |
           ((uint64_t)((z >> 9) + 2048) << 27) | (off4 << 12) | off3;
}

__global__ __launch_bounds__(256) void k_dfs_keys(const int32_t *keys, uint32_t n, uint64_t *dk, uint32_t *idx) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    dk[i] = dfs_key(keys[3u * i], keys[3u * i + 1u], keys[3u * i + 2u]);
    idx[i] = i;
}

// Leaves into sorted order: one wave per leaf, a mask byte per lane.
__global__ __launch_bounds__(256) void k_gather_leaves(const uint32_t *order, uint32_t n, const int32_t *keys,
                                                       const uint8_t *masks, int32_t *keys_out, uint8_t *masks_out) {
    const uint32_t w = blockIdx.x * 4u + (threadIdx.x >> 6), lane = threadIdx.x & 63u;
    if (w >= n) return;
    const uint32_t src = order[w];
    masks_out[(size_t)w * 64u + lane] = masks[(size_t)src * 64u + lane];
    if (lane < 3u) keys_out[(size_t)w * 3u + lane] = keys[(size_t)src * 3u + lane];
}

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

// One slab's leaves in DFS order, kept on the device until every slab is done.
struct SlabOut {
    DevBuf keys, masks;  // n x 3 int32, n x 64 bytes
    uint32_t n = 0;
};

}  // namespace

extern "C" int wx_vdb_add_noise_device(wx_vdb *v, uint64_t seed, const int32_t lo[3], const int32_t hi[3],
                                       float base_freq, float occupancy, wx_stream stream, float *threshold) {
    if (!v || !lo || !hi) return set_error(WX_ERR_INVALID, "wx_vdb_add_noise_device: NULL argument");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return set_error(WX_ERR_DEVICE, "wx_vdb_add_noise_device: no HIP device");
    try {
        std::string err;
        wxnoise::Params p;
        if (!v->tree.noise_params(seed, lo, hi, base_freq, occupancy, p, err))
            return set_error(v->tree.finalized() ? WX_ERR_STATE : WX_ERR_INVALID, "wx_vdb_add_noise_device: " + err);
        LeafBox bx;
        uint64_t n[3];
        for (int a = 0; a < 3; ++a) {
            bx.l0[a] = lo[a] >> 3;
            n[a] = (uint64_t)(((hi[a] - 1) >> 3) - bx.l0[a] + 1);
        }
        bx.ny = (uint32_t)n[1];
        bx.nz = (uint32_t)n[2];
        const uint64_t total = n[0] * n[1] * n[2], plane = n[1] * n[2];
        // slabs of whole x-planes ending on node-4 boundaries (16 leaf planes): each
        // slab's leaves are sorted into DFS order on the device, so within one node-5
        // column in y and z the slabs arrive in DFS order as a whole and the tree
        // needs no sort (elsewhere it sorts them itself)
        if (plane > (1ull << 26))
            return set_error(WX_ERR_UNSUPPORTED, "wx_vdb_add_noise_device: more than 2^26 leaves per x-plane");
        uint64_t planes = 16u * (kSlab / (16u * plane));
        if (!planes) planes = 16u;  // at most 2^30 candidates
        const uint64_t slab = std::min(total, planes * plane);
        // a tree without leaves takes the slabs as they come (every key is new);
        // otherwise they merge like add_leaf_masks
        const bool fresh = v->tree.leaf_count() == 0;
        hipStream_t st = static_cast<hipStream_t>(stream);
        DevBuf counter, keys, masks, todo, skeys, smasks, dk, dk2, idx, idx2, tmp;
        size_t tmp_bytes = 0;
        if (hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, (const uint64_t *)nullptr, (uint64_t *)nullptr,
                                               (const uint32_t *)nullptr, (uint32_t *)nullptr, (uint64_t)slab, 0, 64,
                                               st) != hipSuccess)
            return set_error(WX_ERR_DEVICE, "wx_vdb_add_noise_device: sort set-up failed");
        if (hipMalloc(&counter.p, 8) != hipSuccess || hipMalloc(&keys.p, slab * 12) != hipSuccess ||
            hipMalloc(&todo.p, slab * 4) != hipSuccess ||
            hipMalloc(&masks.p, slab * 64) != hipSuccess || hipMalloc(&skeys.p, slab * 12) != hipSuccess ||
            hipMalloc(&smasks.p, slab * 64) != hipSuccess || hipMalloc(&dk.p, slab * 8) != hipSuccess ||
            hipMalloc(&dk2.p, slab * 8) != hipSuccess || hipMalloc(&idx.p, slab * 4) != hipSuccess ||
            hipMalloc(&idx2.p, slab * 4) != hipSuccess || hipMalloc(&tmp.p, tmp_bytes ? tmp_bytes : 16) != hipSuccess)
            return set_error(WX_ERR_BUDGET, "wx_vdb_add_noise_device: hipMalloc of the slab buffers failed");
        // each slab's sorted leaves stay on the device until the last slab: the tree then
        // grows once, by the exact count (no regrowth, no zero fill), and they come down
        // through pinned staging
        std::vector<std::unique_ptr<SlabOut>> outs;
        uint64_t total_found = 0;
        for (uint64_t p0 = 0; p0 < n[0];) {
            // planes [p0, p1): to the next multiple of 16 leaf planes in absolute x, then
            // `planes` more at a time
            uint64_t p1 = p0 + planes;
            const int64_t ax = (int64_t)bx.l0[0] + (int64_t)p1;
            p1 -= (uint64_t)(((ax % 16) + 16) % 16);
            if (p1 <= p0) p1 = p0 + planes;
            if (p1 > n[0]) p1 = n[0];
            const uint64_t first = p0 * plane;
            const uint64_t count64 = (p1 - p0) * plane;  // at most slab
            if (count64 > 0xFFFFFFFFull) return set_error(WX_ERR_UNSUPPORTED, "wx_vdb_add_noise_device: slab too large");
            const uint32_t count = (uint32_t)count64;
            p0 = p1;
            // counters[0]: output slots, counters[1]: candidates to decide voxel by voxel
            uint32_t cnt[2] = {0, 0};
            uint32_t *dcnt = static_cast<uint32_t *>(counter.p);
            if (hipMemcpyAsync(dcnt, cnt, 8, hipMemcpyHostToDevice, st) != hipSuccess)
                return set_error(WX_ERR_DEVICE, "wx_vdb_add_noise_device: counter reset failed");
            const uint64_t lanes = (count + kPerLane - 1u) / kPerLane;
            hipLaunchKernelGGL(k_noise_classify, dim3((uint32_t)((lanes + 255u) / 256u)), dim3(256), 0, st, p, bx, first,
                               count, dcnt,
                               static_cast<int32_t *>(keys.p), static_cast<uint8_t *>(masks.p),
                               static_cast<uint32_t *>(todo.p), dcnt + 1);
            if (hipGetLastError() != hipSuccess ||
                hipMemcpyAsync(cnt, dcnt, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess)
                return set_error(WX_ERR_DEVICE, "wx_vdb_add_noise_device: kernel failed");
            if (cnt[0] > count || cnt[1] > count - cnt[0])
                return set_error(WX_ERR_DEVICE, "wx_vdb_add_noise_device: slot count out of range");
            uint32_t found = cnt[0];
            if (cnt[1]) {
                const uint32_t waves = (cnt[1] + kPerLane - 1u) / kPerLane;
                hipLaunchKernelGGL(k_noise_voxels, dim3((waves + 3u) / 4u), dim3(256), 0, st, p, bx, first,
                                   static_cast<const uint32_t *>(todo.p), cnt[1], dcnt,
                                   static_cast<int32_t *>(keys.p), static_cast<uint8_t *>(masks.p));
                if (hipGetLastError() != hipSuccess ||
                    hipMemcpyAsync(&found, dcnt, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
                    hipStreamSynchronize(st) != hipSuccess)
                    return set_error(WX_ERR_DEVICE, "wx_vdb_add_noise_device: kernel failed");
                if (found > count) return set_error(WX_ERR_DEVICE, "wx_vdb_add_noise_device: slot count out of range");
            }
            if (!found) continue;
            std::unique_ptr<SlabOut> o(new SlabOut);
            o->n = found;
            if (hipMalloc(&o->keys.p, (size_t)found * 12) != hipSuccess ||
                hipMalloc(&o->masks.p, (size_t)found * 64) != hipSuccess)
                return set_error(WX_ERR_BUDGET, "wx_vdb_add_noise_device: hipMalloc of the slab's leaves failed");
            // the slab's leaves into DFS order (they were appended in no particular order)
            hipLaunchKernelGGL(k_dfs_keys, dim3((found + 255u) / 256u), dim3(256), 0, st,
                               static_cast<const int32_t *>(keys.p), found, static_cast<uint64_t *>(dk.p),
                               static_cast<uint32_t *>(idx.p));
            size_t tb = tmp_bytes;
            if (hipGetLastError() != hipSuccess ||
                hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, static_cast<const uint64_t *>(dk.p),
                                                   static_cast<uint64_t *>(dk2.p), static_cast<const uint32_t *>(idx.p),
                                                   static_cast<uint32_t *>(idx2.p), (uint64_t)found, 0, 64, st) != hipSuccess)
                return set_error(WX_ERR_DEVICE, "wx_vdb_add_noise_device: sort failed");
            hipLaunchKernelGGL(k_gather_leaves, dim3((found + 3u) / 4u), dim3(256), 0, st,
                               static_cast<const uint32_t *>(idx2.p), found, static_cast<const int32_t *>(keys.p),
                               static_cast<const uint8_t *>(masks.p), static_cast<int32_t *>(o->keys.p),
                               static_cast<uint8_t *>(o->masks.p));
            if (hipGetLastError() != hipSuccess) return set_error(WX_ERR_DEVICE, "wx_vdb_add_noise_device: gather failed");
            total_found += found;
            outs.push_back(std::move(o));
        }
        if (hipStreamSynchronize(st) != hipSuccess)
            return set_error(WX_ERR_DEVICE, "wx_vdb_add_noise_device: kernel failed");
        if (total_found) {
            wx::PinnedCopier dl(st);
            if (!dl.ok()) return set_error(WX_ERR_DEVICE, "wx_vdb_add_noise_device: pinned staging allocation failed");
            if (fresh) {
                const size_t base = v->tree.leaf_count();
                int32_t *hk = nullptr;
                uint64_t *hm = nullptr;
                if (!v->tree.grow_new_leaves(total_found, &hk, &hm, err))
                    return set_error(WX_ERR_INVALID, "wx_vdb_add_noise_device: " + err);
                size_t at = 0;
                for (const auto &o : outs) {
                    if (!dl.to_host(hk + 3 * at, o->keys.p, (size_t)o->n * 12) ||
                        !dl.to_host(hm + 8 * at, o->masks.p, (size_t)o->n * 64)) {
                        v->tree.drop_leaves_from(base);  // no uninitialised leaf stays in the tree
                        return set_error(WX_ERR_DEVICE, "wx_vdb_add_noise_device: download failed");
                    }
                    at += o->n;
                }
                // a tree that had no leaves also keeps the masks on the device, in its leaf
                // order, for the device SDF replay (skipped, not an error, without room)
                if (base == 0) {
                    void *all = nullptr;
                    bool kept = hipMalloc(&all, total_found * 64) == hipSuccess;
                    for (size_t k = 0, off = 0; kept && k < outs.size(); off += outs[k]->n, ++k)
                        kept = hipMemcpyAsync(static_cast<char *>(all) + off * 64, outs[k]->masks.p,
                                              (size_t)outs[k]->n * 64, hipMemcpyDeviceToDevice, st) == hipSuccess;
                    kept = hipStreamSynchronize(st) == hipSuccess && kept;
                    int dev = 0;
                    if (kept && hipGetDevice(&dev) == hipSuccess) {
                        v->tree.set_device_masks(std::unique_ptr<wx::DeviceBuffer>(new HipMasks(all, dev)));
                    } else {
                        if (all) (void)hipFree(all);
                        (void)hipGetLastError();  // a failed allocation is not this call's error
                    }
                }
            } else {  // merged like add_leaf_masks
                std::vector<int32_t> hk;
                std::vector<uint64_t> hm;
                for (const auto &o : outs) {
                    hk.resize((size_t)o->n * 3);
                    hm.resize((size_t)o->n * 8);
                    if (!dl.to_host(hk.data(), o->keys.p, (size_t)o->n * 12) ||
                        !dl.to_host(hm.data(), o->masks.p, (size_t)o->n * 64))
                        return set_error(WX_ERR_DEVICE, "wx_vdb_add_noise_device: download failed");
                    if (!v->tree.add_leaf_masks(hk.data(), hm.data(), o->n, err))
                        return set_error(WX_ERR_INVALID, "wx_vdb_add_noise_device: " + err);
                }
            }
        }
        if (threshold) *threshold = p.thr;
        return WX_OK;
    } catch (const std::bad_alloc &) {
        return set_error(WX_ERR_BUDGET, "wx_vdb_add_noise_device: host allocation failed");
    }
}
