// Device generator of the value-noise isosurface leaves: the voxels of
// HostVdb::add_noise, computed by the GPU in slabs of candidate leaves and appended
// to the host tree. The noise, the leaf classes and the threshold are the shared
// definitions of noise.hpp, so both generators produce the same leaves bit for bit.
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#define WX_NOISE_FN __host__ __device__ inline
#include "noise.hpp"
#include "wx_internal.hpp"

using wx::set_error;

namespace {

constexpr uint64_t kSlab = 1ull << 24;  // candidate leaves per launch (1.3 GB of output at most)

struct LeafBox {
    int32_t l0[3];  // first leaf key
    uint32_t ny, nz;
};

// One wave per candidate leaf, c = first + wave, x-major over the leaf box. Lane
// (x, y) = (lane >> 3, lane & 7) decides the voxels (x, y, 0..7), which are bits
// 0..7 of byte y of mask word x: byte `lane` of the leaf's 64-byte mask. A leaf with
// an active voxel takes the next output slot.
__global__ __launch_bounds__(256) void k_noise_leaves(wxnoise::Params p, LeafBox bx, uint64_t first, uint32_t count,
                                                      uint32_t *counter, int32_t *keys, uint8_t *masks) {
    const uint32_t w = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (w >= count) return;  // the whole wave
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t c = first + w;
    const uint64_t r = c / bx.nz;
    const int32_t kx = bx.l0[0] + (int32_t)(r / bx.ny);
    const int32_t ky = bx.l0[1] + (int32_t)(r % bx.ny);
    const int32_t kz = bx.l0[2] + (int32_t)(c % bx.nz);
    const int32_t b[3] = {kx * 8, ky * 8, kz * 8};
    bool inside;
    const int cls = wxnoise::leaf_class(p, b, inside);
    if (cls == 0) return;
    uint32_t byte = 0xFFu;
    if (cls == 2) {
        byte = 0;
        const int32_t x = b[0] + (int32_t)(lane >> 3), y = b[1] + (int32_t)(lane & 7u);
        for (int z = 0; z < 8; ++z)
            if (wxnoise::voxel_active(p, x, y, b[2] + z)) byte |= 1u << z;
    }
    if (__ballot(byte != 0u) == 0) return;
    uint32_t slot = 0;
    if (lane == 0) slot = atomicAdd(counter, 1u);
    slot = __shfl(slot, 0);
    masks[(size_t)slot * 64u + lane] = (uint8_t)byte;
    if (lane < 3u) keys[(size_t)slot * 3u + lane] = lane == 0u ? kx : (lane == 1u ? ky : kz);
}

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
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
        const uint64_t total = n[0] * n[1] * n[2];
        const uint64_t slab = total < kSlab ? total : kSlab;
        // a tree without leaves takes the slabs as they come (every key is new);
        // otherwise they merge like add_leaf_masks
        const bool fresh = v->tree.leaf_count() == 0;
        hipStream_t st = static_cast<hipStream_t>(stream);
        DevBuf counter, keys, masks;
        if (hipMalloc(&counter.p, 4) != hipSuccess || hipMalloc(&keys.p, slab * 12) != hipSuccess ||
            hipMalloc(&masks.p, slab * 64) != hipSuccess)
            return set_error(WX_ERR_BUDGET, "wx_vdb_add_noise_device: hipMalloc of the slab buffers failed");
        std::vector<int32_t> hk;
        std::vector<uint64_t> hm;
        for (uint64_t first = 0; first < total; first += slab) {
            const uint32_t count = (uint32_t)(total - first < slab ? total - first : slab);
            uint32_t found = 0;
            if (hipMemcpyAsync(counter.p, &found, 4, hipMemcpyHostToDevice, st) != hipSuccess)
                return set_error(WX_ERR_DEVICE, "wx_vdb_add_noise_device: counter reset failed");
            hipLaunchKernelGGL(k_noise_leaves, dim3((count + 3u) / 4u), dim3(256), 0, st, p, bx, first, count,
                               static_cast<uint32_t *>(counter.p), static_cast<int32_t *>(keys.p),
                               static_cast<uint8_t *>(masks.p));
            if (hipGetLastError() != hipSuccess ||
                hipMemcpyAsync(&found, counter.p, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess)
                return set_error(WX_ERR_DEVICE, "wx_vdb_add_noise_device: kernel failed");
            if (found > count) return set_error(WX_ERR_DEVICE, "wx_vdb_add_noise_device: slot count out of range");
            if (!found) continue;
            hk.resize((size_t)found * 3);
            hm.resize((size_t)found * 8);
            if (hipMemcpyAsync(hk.data(), keys.p, (size_t)found * 12, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipMemcpyAsync(hm.data(), masks.p, (size_t)found * 64, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess)
                return set_error(WX_ERR_DEVICE, "wx_vdb_add_noise_device: download failed");
            const bool ok = fresh ? v->tree.append_new_leaves(hk.data(), hm.data(), found, err)
                                  : v->tree.add_leaf_masks(hk.data(), hm.data(), found, err);
            if (!ok) return set_error(WX_ERR_INVALID, "wx_vdb_add_noise_device: " + err);
        }
        if (threshold) *threshold = p.thr;
        return WX_OK;
    } catch (const std::bad_alloc &) {
        return set_error(WX_ERR_BUDGET, "wx_vdb_add_noise_device: host allocation failed");
    }
}
