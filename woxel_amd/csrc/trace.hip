// HDDA + grid-SDF traversal kernels for gfx950, and the device-side C ABI.
//
// The per-ray arithmetic restates hdda_ray (src/shaders/raycast.comp.wgsl:84-126)
// operation for operation, in FP32 with contraction off (-ffp-contract=off), IEEE
// division and square root, so results are bit-identical to the scalar oracle.
// The lookups restate get_vdb_leaf_from_{leaf,nothing,node5,node4,node3}
// (raycast.comp.wgsl:360-494) over the packed device layout: one 16-byte record
// per internal-node lookup (child/value bits, popcount base, tile SDF) and one
// 4-bit slot per voxel lookup (0 = active, else SDF), so every tree level costs a
// single gather.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <string>

#include "wx_internal.hpp"

using wx::set_error;

namespace {

struct GridDev {
    const int4 *root;     // n5 x {kx, ky, kz, index}
    const uint4 *n5;      // n5 x 4096 records
    const uint4 *n4;      // n4 x 256 records
    const uint32_t *leaf; // n3 x 64 words
    uint32_t n5count;
    float bound;
};

__device__ __forceinline__ float smin(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float smax(float a, float b) { return a < b ? b : a; }

// f32 -> i32: truncate, saturate, NaN -> 0.
__device__ __forceinline__ int32_t f2i(float x) {
    if (x != x) return 0;
    if (x >= 2147483648.0f) return 2147483647;
    if (x <= -2147483648.0f) return (-2147483647 - 1);
    return (int32_t)x;
}

struct RayOut {
    float px, py, pz;
    uint32_t state;
    uint32_t mask;  // bit c = axis c of the last step
};

// Entry clip (extension of the reference's bound test): a ray starting outside
// [-bound, bound]^3 starts at its slab entry point clamped into the box.
__device__ __forceinline__ bool entry_clip(float ox, float oy, float oz, float dx, float dy, float dz, float ix,
                                           float iy, float iz, float B, float &px, float &py, float &pz) {
    float tn = -__builtin_inff(), tf = __builtin_inff();
    {
        float a = (-B - ox) * ix, b = (B - ox) * ix;
        tn = smax(tn, smin(a, b));
        tf = smin(tf, smax(a, b));
    }
    {
        float a = (-B - oy) * iy, b = (B - oy) * iy;
        tn = smax(tn, smin(a, b));
        tf = smin(tf, smax(a, b));
    }
    {
        float a = (-B - oz) * iz, b = (B - oz) * iz;
        tn = smax(tn, smin(a, b));
        tf = smin(tf, smax(a, b));
    }
    if (!(tn <= tf) || tf < 0.0f) return false;
    px = smin(smax(ox + tn * dx, -B), B);
    py = smin(smax(oy + tn * dy, -B), B);
    pz = smin(smax(oz + tn * dz, -B), B);
    return true;
}

// hdda_ray for one ray; writes the 32-byte hit record.
__device__ __forceinline__ void hdda(const GridDev &g, float ox, float oy, float oz, float dx, float dy, float dz,
                                     uint32_t flags, wx_hit *out) {
    const float sx = dx < 0.0f ? -1.0f : 1.0f;  // sign11 (:70-76)
    const float sy = dy < 0.0f ? -1.0f : 1.0f;
    const float sz = dz < 0.0f ? -1.0f : 1.0f;
    const float ax = smax(0.0f, sx), ay = smax(0.0f, sy), az = smax(0.0f, sz);
    const float ix = 1.0f / dx, iy = 1.0f / dy, iz = 1.0f / dz;
    const float B = g.bound;
    float px = ox, py = oy, pz = oz;
    uint32_t mask = 0;
    uint32_t s01 = 0, s23 = 0, lk = 0;
    uint32_t state = WX_MAXSTEPS;
    // cursor cache (VdbLeaf.parents): node keys and pool indices per level
    uint32_t np = 0;
    int32_t c5x = 0, c5y = 0, c5z = 0, c4x = 0, c4y = 0, c4z = 0, c3x = 0, c3y = 0, c3z = 0;
    uint32_t i5 = 0, i4 = 0, i3 = 0;
    bool run = true;
    if ((flags & WX_TRACE_CLIP) && (B < fabsf(ox) || B < fabsf(oy) || B < fabsf(oz))) {
        if (!entry_clip(ox, oy, oz, dx, dy, dz, ix, iy, iz, B, px, py, pz)) {
            state = WX_OOB;
            px = ox, py = oy, pz = oz;
            run = false;
        }
    }
    for (uint32_t i = 0; run && i < 1000u; ++i) {
        const int32_t x = f2i(floorf(px)), y = f2i(floorf(py)), z = f2i(floorf(pz));
        const int32_t k5x = x >> 12, k5y = y >> 12, k5z = z >> 12;
        const int32_t k4x = x >> 7, k4y = y >> 7, k4z = z >> 7;
        const int32_t k3x = x >> 3, k3y = y >> 3, k3z = z >> 3;
        // get_vdb_leaf_from_leaf (:360-396): deepest cached parent containing pos
        uint32_t lvl;
        if (np == 3 && k3x == c3x && k3y == c3y && k3z == c3z) lvl = 3;
        else if (np >= 2 && k4x == c4x && k4y == c4y && k4z == c4z) lvl = 2;
        else if (np >= 1 && k5x == c5x && k5y == c5y && k5z == c5z) lvl = 1;
        else lvl = 0;
        uint32_t dist = 0;
        bool done = false;
        if (lvl == 0) {  // get_vdb_leaf_from_nothing (:398-413)
            int32_t found = -1;
            for (uint32_t r = 0; r < g.n5count; ++r) {
                const int4 e = g.root[r];
                if (e.x == k5x && e.y == k5y && e.z == k5z) { found = e.w; break; }
            }
            if (found < 0) {
                dist = 1;
                np = 0;
                done = true;
            } else {
                c5x = k5x, c5y = k5y, c5z = k5z;
                i5 = (uint32_t)found;
                np = 1;
                lvl = 1;
            }
        }
        if (!done && lvl == 1) {  // get_vdb_leaf_from_node5 (:415-444)
            lk += 1u;
            const uint32_t off = ((uint32_t)(x >> 7) & 31u) << 10 | ((uint32_t)(y >> 7) & 31u) << 5 | ((uint32_t)(z >> 7) & 31u);
            const uint4 r = g.n5[(size_t)i5 * 4096u + (off >> 3)];
            const uint32_t b = off & 7u;
            np = 1;
            if ((r.x >> (8u + b)) & 1u) {
                dist = 0;
                done = true;
            } else if (!((r.x >> b) & 1u)) {
                dist = ((b < 4u ? r.z : r.w) >> ((b & 3u) * 8u)) & 0xFFu;
                done = true;
            } else {
                i4 = r.y + (uint32_t)__builtin_popcount(r.x & ((1u << b) - 1u));
                c4x = k4x, c4y = k4y, c4z = k4z;
                np = 2;
                lvl = 2;
            }
        }
        if (!done && lvl == 2) {  // get_vdb_leaf_from_node4 (:446-475)
            lk += 1u << 10;
            const uint32_t off = ((uint32_t)(x >> 3) & 15u) << 8 | ((uint32_t)(y >> 3) & 15u) << 4 | ((uint32_t)(z >> 3) & 15u);
            const uint4 r = g.n4[(size_t)i4 * 256u + (off >> 4)];
            const uint32_t b = off & 15u;
            np = 2;
            if ((r.x >> (16u + b)) & 1u) {
                dist = 0;
                done = true;
            } else if (!((r.x >> b) & 1u)) {
                dist = ((b < 8u ? r.z : r.w) >> ((b & 7u) * 4u)) & 15u;
                done = true;
            } else {
                i3 = r.y + (uint32_t)__builtin_popcount(r.x & ((1u << b) - 1u));
                c3x = k3x, c3y = k3y, c3z = k3z;
                np = 3;
                lvl = 3;
            }
        }
        if (!done) {  // get_vdb_leaf_from_node3 (:477-494)
            lk += 1u << 20;
            const uint32_t off = ((uint32_t)x & 7u) << 6 | ((uint32_t)y & 7u) << 3 | ((uint32_t)z & 7u);
            const uint32_t w = g.leaf[(size_t)i3 * 64u + (off >> 3)];
            dist = (w >> ((off & 7u) * 4u)) & 15u;
            np = 3;
        }
        if (np < 2) s01 += np == 0 ? 1u : 0x10000u;
        else s23 += np == 2 ? 1u : 0x10000u;
        if (dist == 0u) { state = WX_HIT; break; }                               // :96-98
        if (B < fabsf(px) || B < fabsf(py) || B < fabsf(pz)) { state = WX_OOB; break; }  // :101-103
        const float sc = np == 3 ? 1.0f : (np == 2 ? 8.0f : (np == 1 ? 128.0f : 4096.0f));
        const float size = (float)dist * sc;                                      // :105-112
        const float mx = px - size * floorf(px / size);                           // :78-80
        const float my = py - size * floorf(py / size);
        const float mz = pz - size * floorf(pz / size);
        const float tx = ix * (size * ax - mx);                                   // :114
        const float ty = iy * (size * ay - my);
        const float tz = iz * (size * az - mz);
        const float tm = smin(smin(tx, ty), tz);                                  // :116
        px = px + tm * dx;
        py = py + tm * dy;
        pz = pz + tm * dz;
        const uint32_t mxb = (tx <= ty) & (tx <= tz);                             // :118-120
        const uint32_t myb = (ty <= tz) & (ty <= tx);
        const uint32_t mzb = (tz <= tx) & (tz <= ty);
        mask = mxb | (myb << 1) | (mzb << 2);
        px = px + (4e-4f * sx) * (float)mxb;                                      // :122
        py = py + (4e-4f * sy) * (float)myb;
        pz = pz + (4e-4f * sz) * (float)mzb;
    }
    float t = (px - ox) * dx;
    t = t + (py - oy) * dy;
    t = t + (pz - oz) * dz;
    const int32_t nx = (mask & 1u) ? -(int32_t)sx : 0;
    const int32_t ny = (mask & 2u) ? -(int32_t)sy : 0;
    const int32_t nz = (mask & 4u) ? -(int32_t)sz : 0;
    uint4 w0, w1;
    w0.x = (uint32_t)f2i(floorf(px));
    w0.y = (uint32_t)f2i(floorf(py));
    w0.z = (uint32_t)f2i(floorf(pz));
    w0.w = __float_as_uint(t);
    w1.x = s01;
    w1.y = s23;
    w1.z = state | ((uint32_t)(nx & 0xFF) << 8) | ((uint32_t)(ny & 0xFF) << 16) | ((uint32_t)(nz & 0xFF) << 24);
    w1.w = lk;
    uint4 *o = reinterpret_cast<uint4 *>(out);
    o[0] = w0;
    o[1] = w1;
}

struct CamDev {
    float ex, ey, ez;
    float ux, uy, uz;
    float mx, my, mz;
    float wx, wy, wz;
};

// cp_main (raycast.comp.wgsl:60-68): one primary ray per pixel. A 256-lane block
// covers a 16x16 pixel tile; each wave64 covers an 8x8 sub-tile.
__global__ __launch_bounds__(256) void k_trace_camera(GridDev g, CamDev c, uint32_t w, uint32_t h, float off,
                                                      uint32_t flags, wx_hit *out) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t px = blockIdx.x * 16u + (wave & 1u) * 8u + (lane & 7u);
    const uint32_t py = blockIdx.y * 16u + (wave >> 1) * 8u + (lane >> 3);
    if (px >= w || py >= h) return;
    const float fx = (float)px + off, fy = (float)py + off;
    const float vx = (fx * c.ux + fy * c.mx) + c.wx;
    const float vy = (fx * c.uy + fy * c.my) + c.wy;
    const float vz = (fx * c.uz + fy * c.mz) + c.wz;
    const float len = sqrtf((vx * vx + vy * vy) + vz * vz);
    hdda(g, c.ex, c.ey, c.ez, vx / len, vy / len, vz / len, flags, out + (size_t)py * w + px);
}

__global__ __launch_bounds__(256) void k_trace_rays(GridDev g, const wx_ray *rays, uint32_t n, uint32_t flags,
                                                    wx_hit *out) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const float *r = reinterpret_cast<const float *>(rays + i);
    hdda(g, r[0], r[1], r[2], r[3], r[4], r[5], flags, out + i);
}

int hip_fail(hipError_t e, const char *what) {
    return set_error(WX_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace

struct wx_grid {
    GridDev dev;
    void *mem[4] = {nullptr, nullptr, nullptr, nullptr};
    uint64_t bytes = 0;
    int device = 0;
};

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
    if ((d->n5 && !d->root) || (d->n5 && !d->n5_records) || (d->n4 && !d->n4_records) || (d->n3 && !d->leaves))
        return set_error(WX_ERR_INVALID, "wx_grid_create: NULL pool with a non-zero count");
    if (!(d->bound > 0.0f)) return set_error(WX_ERR_INVALID, "wx_grid_create: bound must be positive");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return set_error(WX_ERR_DEVICE, "wx_grid_create: no HIP device");
    wx_grid *g = new (std::nothrow) wx_grid();
    if (!g) return set_error(WX_ERR_BUDGET, "wx_grid_create: host allocation failed");
    hipGetDevice(&g->device);
    const size_t sz[4] = {(size_t)d->n5 * 16u, (size_t)d->n5 * 4096u * 16u, (size_t)d->n4 * 256u * 16u,
                          (size_t)d->n3 * 64u * 4u};
    const void *src[4] = {d->root, d->n5_records, d->n4_records, d->leaves};
    for (int k = 0; k < 4; ++k) {
        size_t n = sz[k] ? sz[k] : 16;
        hipError_t e = hipMalloc(&g->mem[k], n);
        if (e != hipSuccess) {
            for (int j = 0; j < k; ++j) hipFree(g->mem[j]);
            delete g;
            return set_error(e == hipErrorOutOfMemory ? WX_ERR_BUDGET : WX_ERR_DEVICE,
                             std::string("wx_grid_create: hipMalloc: ") + hipGetErrorString(e));
        }
        g->bytes += n;
        if (sz[k]) {
            e = hipMemcpy(g->mem[k], src[k], sz[k], hipMemcpyHostToDevice);
            if (e != hipSuccess) {
                for (int j = 0; j <= k; ++j) hipFree(g->mem[j]);
                delete g;
                return hip_fail(e, "wx_grid_create: hipMemcpy");
            }
        }
    }
    g->dev.root = static_cast<const int4 *>(g->mem[0]);
    g->dev.n5 = static_cast<const uint4 *>(g->mem[1]);
    g->dev.n4 = static_cast<const uint4 *>(g->mem[2]);
    g->dev.leaf = static_cast<const uint32_t *>(g->mem[3]);
    g->dev.n5count = d->n5;
    g->dev.bound = d->bound;
    *out = g;
    return WX_OK;
}

int wx_grid_destroy(wx_grid *g) {
    if (!g) return WX_OK;
    for (int k = 0; k < 4; ++k)
        if (g->mem[k]) hipFree(g->mem[k]);
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
    if (w == 0 || h == 0) return WX_OK;
    if (w > 65536u * 16u || h > 65536u * 16u) return set_error(WX_ERR_INVALID, "wx_trace_camera: image too large");
    CamDev c;
    c.ex = cam->eye[0], c.ey = cam->eye[1], c.ez = cam->eye[2];
    c.ux = cam->u[0], c.uy = cam->u[1], c.uz = cam->u[2];
    c.mx = cam->mv[0], c.my = cam->mv[1], c.mz = cam->mv[2];
    c.wx = cam->wp[0], c.wy = cam->wp[1], c.wz = cam->wp[2];
    dim3 grid((w + 15u) / 16u, (h + 15u) / 16u, 1);
    hipLaunchKernelGGL(k_trace_camera, grid, dim3(256), 0, (hipStream_t)stream, g->dev, c, w, h, pixel_offset, flags,
                       out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "wx_trace_camera: launch");
    return WX_OK;
}

int wx_trace_rays(wx_grid *g, const wx_ray *rays, uint32_t n, uint32_t flags, wx_hit *out, wx_stream stream) {
    if (!g || ((!rays || !out) && n)) return set_error(WX_ERR_INVALID, "wx_trace_rays: NULL argument");
    if (n == 0) return WX_OK;
    dim3 grid((n + 255u) / 256u, 1, 1);
    hipLaunchKernelGGL(k_trace_rays, grid, dim3(256), 0, (hipStream_t)stream, g->dev, rays, n, flags, out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "wx_trace_rays: launch");
    return WX_OK;
}

}  // extern "C"
