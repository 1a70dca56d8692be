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

__device__ __forceinline__ bool bit(const uint64_t *m, uint32_t b) { return (m[b >> 6] >> (b & 63)) & 1ull; }

// One workgroup (one wave) per node of one schedule group.
template <int LOG>
__global__ __launch_bounds__(64) void k_sdf_nodes(const uint64_t *solid, uint8_t *val, const int32_t *nb,
                                                  const uint32_t *group, const uint16_t *sweep_slots,
                                                  const uint32_t *sweep_start, int forward) {
    constexpr uint32_t D = 1u << LOG, S = D * D * D, W = S / 64, T = 7 * (D - 1) + 1;
    __shared__ uint8_t v_lds[S];
    __shared__ uint64_t s_lds[W];
    const uint32_t n = group[blockIdx.x];
    const uint64_t *sn = solid + (size_t)n * W;
    for (uint32_t i = threadIdx.x; i < W; i += 64u) s_lds[i] = sn[i];
    for (uint32_t i = threadIdx.x; i < S; i += 64u) v_lds[i] = val[(size_t)n * S + i];
    __syncthreads();
    const int32_t *nbn = nb + (size_t)n * 27;
    const int sgn = forward ? -1 : 1;  // the pass reads (sgn * 1, ...) offsets
    for (uint32_t tt = 0; tt < T; ++tt) {
        const uint32_t t = forward ? tt : T - 1u - tt;
        for (uint32_t i = sweep_start[t] + threadIdx.x; i < sweep_start[t + 1]; i += 64u) {
            const uint32_t s = sweep_slots[i];
            if (bit(s_lds, s)) continue;
            uint32_t v = v_lds[s];
            if (v == 1u) continue;
            const int x = (int)(s >> (2 * LOG)), y = (int)((s >> LOG) & (D - 1)), z = (int)(s & (D - 1));
            // the 13 neighbours: (sgn, dy, dz), (0, sgn, dz), (0, 0, sgn) for dy, dz in -1..1
            for (int k = 0; k < 13 && v != 1u; ++k) {
                int dx, dy, dz;
                if (k < 9) dx = sgn, dy = k / 3 - 1, dz = k % 3 - 1;
                else if (k < 12) dx = 0, dy = sgn, dz = k - 10;
                else dx = 0, dy = 0, dz = sgn;
                int cx = x + dx, cy = y + dy, cz = z + dz;
                const int ox = cx < 0 ? -1 : (cx >= (int)D ? 1 : 0);
                const int oy = cy < 0 ? -1 : (cy >= (int)D ? 1 : 0);
                const int oz = cz < 0 ? -1 : (cz >= (int)D ? 1 : 0);
                cx &= D - 1, cy &= D - 1, cz &= D - 1;
                const uint32_t c = ((uint32_t)cx << (2 * LOG)) | ((uint32_t)cy << LOG) | (uint32_t)cz;
                uint32_t w;
                if ((ox | oy | oz) == 0) {  // in this node
                    if (bit(s_lds, c)) {
                        v = 1u;
                        break;
                    }
                    w = v_lds[c];
                } else {  // get_voxel in a neighbouring node: same-level empty slot, else 1
                    const int32_t m = nbn[(ox + 1) * 9 + (oy + 1) * 3 + (oz + 1)];
                    if (m < 0 || bit(solid + (size_t)m * W, c)) {
                        v = 1u;
                        break;
                    }
                    w = val[(size_t)m * S + c];
                }
                w = w == kUnknown ? (uint32_t)kUnknown : w + 1u;
                v = w < v ? w : v;
            }
            v_lds[s] = (uint8_t)v;
        }
        __syncthreads();
    }
    for (uint32_t i = threadIdx.x; i < S; i += 64u) val[(size_t)n * S + i] = v_lds[i];
}

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

struct Ctx {
    hipStream_t stream;
    int status = WX_OK;
    uint64_t launches = 0;
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

template <int LOG>
bool run_level(wx::SdfLevelJob &J, Ctx &c, std::string &err) {
    constexpr size_t S = (size_t)1 << (3 * LOG), W = S / 64;
    if (!J.nodes) return true;
    const Sweep sw = make_sweep(LOG);
    DevBuf solid, val, nb, ord0, ord1, sslots, sstart;
    if (!upload(solid, J.solid, J.nodes * W, c.stream, err) || !upload(val, J.val, J.nodes * S, c.stream, err) ||
        !upload(nb, J.nb.data(), J.nb.size(), c.stream, err) ||
        !upload(ord1, J.order[1].data(), J.order[1].size(), c.stream, err) ||
        !upload(ord0, J.order[0].data(), J.order[0].size(), c.stream, err) ||
        !upload(sslots, sw.slots.data(), sw.slots.size(), c.stream, err) ||
        !upload(sstart, sw.start.data(), sw.start.size(), c.stream, err))
        return false;
    for (int fw = 1; fw >= 0; --fw) {  // forward pass, then backward
        const std::vector<uint32_t> &st = J.start[fw];
        const uint32_t *ord = static_cast<const uint32_t *>(fw ? ord1.p : ord0.p);
        for (size_t g = 0; g + 1 < st.size(); ++g) {
            const uint32_t cnt = st[g + 1] - st[g];
            if (!cnt) continue;
            hipLaunchKernelGGL(k_sdf_nodes<LOG>, dim3(cnt), dim3(64), 0, c.stream,
                               static_cast<const uint64_t *>(solid.p), static_cast<uint8_t *>(val.p),
                               static_cast<const int32_t *>(nb.p), ord + st[g],
                               static_cast<const uint16_t *>(sslots.p), static_cast<const uint32_t *>(sstart.p), fw);
            ++c.launches;
        }
    }
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(J.val, val.p, J.nodes * S, hipMemcpyDeviceToHost, c.stream) != hipSuccess ||
        hipStreamSynchronize(c.stream) != hipSuccess) {
        err = "device SDF: kernel or download failed";
        c.status = WX_ERR_DEVICE;
        return false;
    }
    return true;
}

bool run_device(wx::SdfLevelJob *jobs, void *ctx, std::string &err) {
    Ctx &c = *static_cast<Ctx *>(ctx);
    return run_level<5>(jobs[0], c, err) && run_level<4>(jobs[1], c, err) && run_level<3>(jobs[2], c, err);
}

}  // namespace

extern "C" int wx_vdb_finalize_device(wx_vdb *v, int threads, wx_stream stream, uint64_t *launches) {
    if (!v) return set_error(WX_ERR_INVALID, "wx_vdb_finalize_device: NULL tree");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        return set_error(WX_ERR_DEVICE, "wx_vdb_finalize_device: no HIP device");
    try {
        Ctx c;
        c.stream = static_cast<hipStream_t>(stream);
        std::string err;
        if (!v->tree.finalize_with(threads < 1 ? 1 : threads, run_device, &c, err))
            return set_error(c.status != WX_OK ? c.status : WX_ERR_UNSUPPORTED, "wx_vdb_finalize_device: " + err);
        if (launches) *launches = c.launches;
        return WX_OK;
    } catch (const std::bad_alloc &) {
        return set_error(WX_ERR_BUDGET, "wx_vdb_finalize_device: host allocation failed");
    }
}
