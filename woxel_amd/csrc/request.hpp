// Request records and hit-record writes shared by the traversal kernels that take
// requests (the batched, persistent and serving kernels).
#pragma once

#include "lane.hpp"

namespace wxk {

// Device copy of a request (wx_request with the host-only fields dropped).
struct ReqDev {
    uint32_t kind, grid, w, h;
    uint32_t flags, tiles_x;
    float off;
    uint32_t pad0;
    float ex, ey, ez, ux, uy, uz, mx, my, mz, wx, wy, wz;
    float lx, ly, lz, pad1;
    const wx_ray *rays;
    wx_hit *out;
    wx_hit *shadow;
    uint64_t pad2;
};
static_assert(sizeof(ReqDev) == 128, "ReqDev layout");

// The 32-byte hit record (wx_hit) of a finished lane as two 16-byte words; (ox, oy,
// oz) is the ray origin before any clip.
__device__ __forceinline__ void hit_words(const RayLane &r, float ox, float oy, float oz, uint4 &w0, uint4 &w1) {
    float t = (r.px - ox) * r.dx;
    t = t + (r.py - oy) * r.dy;
    t = t + (r.pz - oz) * r.dz;
    const int32_t nx = r.ax != 0.0f ? (r.dx < 0.0f ? 1 : -1) : 0;
    const int32_t ny = r.ay != 0.0f ? (r.dy < 0.0f ? 1 : -1) : 0;
    const int32_t nz = r.az != 0.0f ? (r.dz < 0.0f ? 1 : -1) : 0;
    w0.x = (uint32_t)f2i(floorf(r.px));
    w0.y = (uint32_t)f2i(floorf(r.py));
    w0.z = (uint32_t)f2i(floorf(r.pz));
    w0.w = __float_as_uint(t);
    w1.x = r.s01;
    w1.y = r.s23;
    w1.z = lane_state(r) | ((uint32_t)(nx & 0xFF) << 8) | ((uint32_t)(ny & 0xFF) << 16) | ((uint32_t)(nz & 0xFF) << 24);
    w1.w = lookup_word(r);
}

__device__ __forceinline__ void lane_write(const RayLane &r, float ox, float oy, float oz, wx_hit *out) {
    uint4 w0, w1;
    hit_words(r, ox, oy, oz, w0, w1);
    uint4 *o = reinterpret_cast<uint4 *>(out);
    o[0] = w0;
    o[1] = w1;
}

__device__ __forceinline__ void write_no_ray(wx_hit *out) {
    uint4 *o = reinterpret_cast<uint4 *>(out);
    o[0] = make_uint4(0, 0, 0, 0);
    o[1] = make_uint4(0, 0, WX_NO_RAY, 0);
}

// cp_main (raycast.comp.wgsl:62-63): dir = normalize(p.x*u + p.y*mv + wp), p = pixel + offset.
__device__ __forceinline__ void camera_dir(const ReqDev &q, uint32_t px, uint32_t py, float &dx, float &dy,
                                           float &dz) {
    camera_ray(&q.ux, &q.mx, &q.wx, q.off, px, py, dx, dy, dz);
}

// Shadow ray from a hit (raycast.comp.wgsl:188-190) toward a point light:
// origin = p - (4e-2 * step) * mask, dir = normalize(light - origin).
__device__ __forceinline__ void shadow_origin(const RayLane &r, float &ox, float &oy, float &oz) {
    const float sx = r.dx < 0.0f ? -1.0f : 1.0f;
    const float sy = r.dy < 0.0f ? -1.0f : 1.0f;
    const float sz = r.dz < 0.0f ? -1.0f : 1.0f;
    ox = r.px - (4e-2f * sx) * (r.ax != 0.0f ? 1.0f : 0.0f);
    oy = r.py - (4e-2f * sy) * (r.ay != 0.0f ? 1.0f : 0.0f);
    oz = r.pz - (4e-2f * sz) * (r.az != 0.0f ? 1.0f : 0.0f);
}

__device__ __forceinline__ void start_shadow(RayLane &r, const ReqDev &q, float B, float ox, float oy, float oz) {
    const float vx = q.lx - ox, vy = q.ly - oy, vz = q.lz - oz;
    const float len = sqrtf((vx * vx + vy * vy) + vz * vz);
    lane_start(r, ox, oy, oz, vx / len, vy / len, vz / len, B, (q.flags & WX_TRACE_CLIP) != 0u);
}

// Camera slot -> pixel: 64 consecutive slots are one 8x8 pixel tile.
__device__ __forceinline__ bool camera_pixel(const ReqDev &q, uint32_t s, uint32_t &px, uint32_t &py) {
    const uint32_t tile = s >> 6, l = s & 63u;
    px = (tile % q.tiles_x) * 8u + (l & 7u);
    py = (tile / q.tiles_x) * 8u + (l >> 3);
    return px < q.w && py < q.h;
}

}  // namespace wxk
