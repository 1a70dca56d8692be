// Shading stage: the colour of each pixel, as the reference's compute pass
// writes it into its rgba8 texture (cp_main + ray_trace, reflect_ray2,
// reflect_ray1; src/shaders/raycast.comp.wgsl:60-68, 144-342), for render modes
// 0 gray, 1 rgb, 2 ray length, 3 diffuse, 4 glossy (RenderMode,
// src/render/egui_dev.rs:12-18).
//
// One lane per pixel. A pixel's rays -- the primary, the shadow ray toward the
// sun, and in glossy mode two reflection bounces with their own shadow rays (up
// to six traversals) -- run one after the other in the lane, through a single
// traversal loop (lane_run): lanes at different stages of their chain share
// the loop's instructions, and the rays of one pass of the loop all start
// together. The colour arithmetic restates the shader's in FP32
// with contraction off: dot = (x + y) + z, normalize = v / sqrt(dot(v, v)),
// mix(a, b, t) = a * (1 - t) + b * t, max(a, b) = a < b ? b : a,
// min(a, b) = b < a ? b : a, the float remainder as fmodf. The rgba8 store
// clamps to [0, 1] and rounds c * 255 to nearest even.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>

#include "lane.hpp"
#include "wx_internal.hpp"

using wx::set_error;

namespace {

using namespace wxk;

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 mul(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ V3 scale(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ V3 adds(V3 a, float s) { return v3(a.x + s, a.y + s, a.z + s); }
__device__ __forceinline__ float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ V3 normalize(V3 a) {
    const float l = sqrtf(dot(a, a));
    return v3(a.x / l, a.y / l, a.z / l);
}
__device__ __forceinline__ V3 mix(V3 a, V3 b, float t) {
    const float u = 1.0f - t;
    return v3(a.x * u + b.x * t, a.y * u + b.y * t, a.z * u + b.z * t);
}

// material constants (raycast.comp.wgsl:144-150)
constexpr float kD = 0.7f, kA = 0.3f, kRefl = 0.9f, kWall = 0.1f;
__device__ __forceinline__ V3 base_color() { return v3(0.4f, 0.2f, 0.2f); }
__device__ __forceinline__ V3 ambient_color() { return v3(0.4f, 0.4f, 0.3f); }

struct ShadeDev {
    GridDev g;
    float cam[12];  // eye, u, mv, wp (xyz each)
    float off;
    uint32_t w, h, flags;
    uint32_t show3, show4, show5, pad;
    float sdx, sdy, sdz, sa;  // sun_dir, sun intensity (sun_color.a)
    float scr, scg, scb, pad2;
    void *out;
    uint64_t pad3;
};

// Result of one hdda_ray call: state, p, mask (as 0/1 floats), the ray's step
// and direction, the iteration of return.
struct Ray {
    uint32_t state;
    V3 p, m, step, d;
    float i;
};

__device__ __forceinline__ V3 sign11(V3 d) {
    return v3(d.x < 0.0f ? -1.0f : 1.0f, d.y < 0.0f ? -1.0f : 1.0f, d.z < 0.0f ? -1.0f : 1.0f);
}

// N = normalize(-step * mask)
__device__ __forceinline__ V3 normal_of(const Ray &h) { return normalize(mul(v3(-h.step.x, -h.step.y, -h.step.z), h.m)); }

// the origin of secondary rays: hit.p - 4e-2 * step * mask (avoid self-intersection)
__device__ __forceinline__ V3 off_origin(const Ray &h) { return sub(h.p, mul(scale(h.step, 4e-2f), h.m)); }

// walls seen by an escaping reflection (reflect_ray2/1, raycast.comp.wgsl:262-274, 322-334)
__device__ __forceinline__ V3 walls(V3 N) {
    const V3 Np = v3(smax(0.0f, N.x), smax(0.0f, N.y), smax(0.0f, N.z));
    const V3 Nn = v3(-smin(0.0f, N.x), -smin(0.0f, N.y), -smin(0.0f, N.z));
    V3 c = scale(v3(kWall, 0.0f, 0.0f), Np.x);
    c = add(c, scale(v3(0.0f, kWall, 0.0f), Np.y));
    c = add(c, scale(v3(0.0f, 0.0f, kWall), Np.z));
    c = add(c, scale(v3(kWall, kWall, 0.0f), Nn.x));
    c = add(c, scale(v3(0.0f, kWall, kWall), Nn.y));
    c = add(c, scale(v3(kWall, 0.0f, kWall), Nn.z));
    return c;
}

// walls seen by an escaping primary in glossy mode, darkening with height (:220-231)
__device__ __forceinline__ V3 walls_primary(V3 N, float py) {
    const V3 Np = v3(smax(0.0f, N.x), smax(0.0f, N.y), smax(0.0f, N.z));
    const V3 Nn = v3(-smin(0.0f, N.x), -smin(0.0f, N.y), -smin(0.0f, N.z));
    const float t = py / 4096.0f;
    const float w1 = kWall * 0.1f;
    V3 c = scale(mix(v3(kWall, 0.0f, 0.0f), v3(w1, 0.0f, 0.0f), t), Np.x);
    c = add(c, scale(v3(0.0f, kWall, 0.0f), Np.y));
    c = add(c, scale(mix(v3(0.0f, 0.0f, kWall), v3(0.0f, 0.0f, w1), t), Np.z));
    c = add(c, scale(mix(v3(kWall, kWall, 0.0f), v3(w1, w1, 0.0f), t), Nn.x));
    c = add(c, scale(v3(0.0f, kWall, kWall), Nn.y));
    c = add(c, scale(mix(v3(kWall, 0.0f, kWall), v3(w1, 0.0f, w1), t), Nn.z));
    return c;
}

// grid overlay on hits at node boundaries (:155-164)
__device__ __forceinline__ V3 grid_overlay(const ShadeDev &a, V3 p) {
    const float fx = floorf(p.x), fy = floorf(p.y), fz = floorf(p.z);
    auto on = [&](float s) { return fmodf(fx, s) == 0.0f || fmodf(fy, s) == 0.0f || fmodf(fz, s) == 0.0f; };
    if (a.show5 == 1u && on(4096.0f)) return v3(-0.3f, -0.3f, 1.0f);
    if (a.show4 == 1u && on(128.0f)) return v3(0.6f, -0.2f, -0.2f);
    if (a.show3 == 1u && on(8.0f)) return v3(-0.1f, 0.5f, 0.3f);
    return v3(0.0f, 0.0f, 0.0f);
}

__device__ __forceinline__ float mask_dot(V3 m, float a, float b, float c) {
    return dot(mul(m, v3(a, b, c)), v3(1.0f, 1.0f, 1.0f));
}

// sun-lit material colour of a glossy hit: BASE + I * sun (x 0.05 in shadow)
__device__ __forceinline__ V3 lit(const ShadeDev &a, float I, bool shadowed) {
    const V3 sc = v3(a.scr, a.scg, a.scb);
    return shadowed ? add(base_color(), scale(scale(sc, I), 0.05f)) : add(base_color(), scale(sc, I));
}

__device__ __forceinline__ float glossy_I(const ShadeDev &a, V3 N) {
    // s.sun_color.a * k_d * dot(-s.sun_dir, N), then max(0, I)
    return smax(0.0f, (a.sa * kD) * dot(v3(-a.sdx, -a.sdy, -a.sdz), N));
}

__device__ __forceinline__ uint32_t unorm8(float c) {
    const float s = c > 0.0f ? (c < 1.0f ? c : 1.0f) : 0.0f;
    return (uint32_t)__builtin_rintf(s * 255.0f);
}

template <uint32_t MODE>
__global__ __launch_bounds__(128) void k_render(ShadeDev a) {
    __shared__ float2 s_tab[kTab];
    __shared__ int4 s_root_buf[kRootLds];
    const bool root_lds = stage_root(a.g, s_root_buf);
    init_tab(s_tab);
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t px = blockIdx.x * 16u + wave * 8u + (lane & 7u);  // 16x8-pixel workgroups
    const uint32_t py = blockIdx.y * 8u + (lane >> 3);
    if (px >= a.w || py >= a.h) return;
    const bool clip = (a.flags & WX_TRACE_CLIP) != 0u;
    const V3 eye = v3(a.cam[0], a.cam[1], a.cam[2]);
    V3 dir;
    camera_ray(a.cam + 3, a.cam + 6, a.cam + 9, a.off, px, py, dir.x, dir.y, dir.z);
    const V3 sun = v3(-a.sdx, -a.sdy, -a.sdz);

    // The chain of rays of this pixel. bounce 0 = primary, 1 = reflect_ray2,
    // 2 = reflect_ray1; each may be followed by its shadow ray toward -sun_dir.
    V3 src = eye, d = dir;
    bool shadow_phase = false;
    uint32_t bounce = 0;
    Ray h;             // the main ray of the current bounce
    V3 m0, m1, last;   // material colours of bounces 0 and 1; colour of the last bounce
    float I = 0.0f;    // LN (diffuse) or I (glossy) of the current hit
    m0 = m1 = last = v3(0.0f, 0.0f, 0.0f);
    for (;;) {
        RayLane r;
        lane_start(r, src.x, src.y, src.z, d.x, d.y, d.z, a.g.bound, clip);
        lane_run(a.g, r, s_tab, nullptr, s_root_buf, root_lds);
        bool shadowed = false;
        if (shadow_phase) {
            shadowed = lane_state(r) == WX_HIT;
        } else {
            h.state = lane_state(r);
            h.p = v3(r.px, r.py, r.pz);
            h.m = v3(r.ax != 0.0f ? 1.0f : 0.0f, r.ay != 0.0f ? 1.0f : 0.0f, r.az != 0.0f ? 1.0f : 0.0f);
            h.step = sign11(d);
            h.d = d;
            const uint32_t it = lane_iterations(r);
            h.i = (float)(it ? it - 1u : 0u);
            if (MODE <= 2u || h.state != WX_HIT) break;
            const V3 N = normal_of(h);
            // diffuse: LN = max(0, sun_color.a * dot(-sun_dir, N)); glossy: the same
            // with k_d (raycast.comp.wgsl:172-173, 189-191); the shadow ray is cast
            // only when the surface is lit
            I = MODE == 3u ? smax(0.0f, a.sa * dot(sun, N)) : glossy_I(a, N);
            if (I != 0.0f) {
                src = off_origin(h);
                d = sun;
                shadow_phase = true;
                continue;
            }
        }
        shadow_phase = false;
        if (MODE == 3u) {
            const V3 Ia = mul(scale(ambient_color(), kA), base_color());
            const V3 Id = scale(mul(scale(v3(a.scr, a.scg, a.scb), kD), base_color()), I);
            last = shadowed ? Ia : add(Ia, Id);
            break;
        }
        const V3 mc = lit(a, I, shadowed);
        if (bounce == 2u) {
            last = mc;
            break;
        }
        if (bounce == 0u) m0 = mc;
        else m1 = mc;
        // reflect about N, from the offset origin (:196-202)
        const V3 N = normal_of(h);
        const V3 rd = normalize(sub(h.d, scale(scale(N, 2.0f), dot(h.d, N))));
        src = off_origin(h);
        d = rd;
        ++bounce;
    }

    V3 c;
    if (MODE <= 2u) {
        if (h.state == WX_HIT) {
            const V3 grid = grid_overlay(a, h.p);
            if (MODE == 0u) c = adds(grid, mask_dot(h.m, 0.2f, 0.2f, 0.3f));
            else if (MODE == 1u) c = add(adds(grid, 0.1f), mul(h.m, v3(0.4f, 0.4f, 0.4f)));
            else c = add(grid, mix(v3(0.72f, 1.0f, 0.99f), v3(1.0f, 0.0f, 0.0f), h.i / 200.0f));
        } else if (h.state == WX_OOB) {
            if (MODE == 2u)
                c = adds(mix(v3(0.72f, 1.0f, 0.99f), v3(1.0f, 0.0f, 0.0f), h.i / 200.0f),
                         mask_dot(h.m, 0.04f, 0.08f, 0.12f));
            else c = adds(v3(0.0f, 0.0f, 0.0f), mask_dot(h.m, 0.01f, 0.02f, 0.03f));
        } else {
            c = dir;
        }
    } else if (MODE == 3u) {
        if (h.state == WX_HIT) c = last;
        else if (h.state == WX_OOB) c = adds(v3(0.0f, 0.0f, 0.0f), mask_dot(h.m, 0.01f, 0.02f, 0.03f));
        else c = dir;
    } else {
        // colour of the last bounce, then folded back: mix(mcol, colour(next), 0.9)
        V3 col;
        if (h.state == WX_HIT) col = last;
        else if (h.state == WX_OOB) col = bounce == 0u ? walls_primary(normal_of(h), h.p.y) : walls(normal_of(h));
        else col = h.d;
        if (bounce == 2u) col = mix(m1, col, kRefl);
        if (bounce >= 1u) col = mix(m0, col, kRefl);
        c = col;
    }

    const uint32_t pix = py * a.w + px;
    if (a.flags & WX_RENDER_F32) {
        reinterpret_cast<float4 *>(a.out)[pix] = make_float4(c.x, c.y, c.z, 1.0f);
    } else {
        reinterpret_cast<uint32_t *>(a.out)[pix] =
            unorm8(c.x) | (unorm8(c.y) << 8) | (unorm8(c.z) << 16) | (255u << 24);
    }
}

}  // namespace

extern "C" int wx_render_camera(wx_grid *g, const wx_camera *cam, uint32_t w, uint32_t h, float pixel_offset,
                                const wx_shade *s, uint32_t flags, void *out, wx_stream stream) {
    if (!g || !cam || !s || (!out && w && h)) return set_error(WX_ERR_INVALID, "wx_render_camera: NULL argument");
    if (flags & ~(WX_TRACE_CLIP | WX_RENDER_F32)) return set_error(WX_ERR_INVALID, "wx_render_camera: unknown flags");
    if (s->mode > 4u) return set_error(WX_ERR_INVALID, "wx_render_camera: render mode must be 0..4");
    if (w > 65536u || h > 65536u) return set_error(WX_ERR_INVALID, "wx_render_camera: image too large");
    if (!w || !h) return WX_OK;
    ShadeDev a;
    std::memset(&a, 0, sizeof a);
    a.g = g->dev;
    for (int c = 0; c < 3; ++c) {
        a.cam[c] = cam->eye[c];
        a.cam[3 + c] = cam->u[c];
        a.cam[6 + c] = cam->mv[c];
        a.cam[9 + c] = cam->wp[c];
    }
    a.off = pixel_offset;
    a.w = w, a.h = h, a.flags = flags;
    a.show3 = s->show_345[0], a.show4 = s->show_345[1], a.show5 = s->show_345[2];
    a.sdx = s->sun_dir[0], a.sdy = s->sun_dir[1], a.sdz = s->sun_dir[2], a.sa = s->sun_color[3];
    a.scr = s->sun_color[0], a.scg = s->sun_color[1], a.scb = s->sun_color[2];
    a.out = out;
    const dim3 grid((w + 15u) / 16u, (h + 7u) / 8u, 1);
    hipStream_t st = static_cast<hipStream_t>(stream);
    switch (s->mode) {
        case 0: hipLaunchKernelGGL(k_render<0>, grid, dim3(128), 0, st, a); break;
        case 1: hipLaunchKernelGGL(k_render<1>, grid, dim3(128), 0, st, a); break;
        case 2: hipLaunchKernelGGL(k_render<2>, grid, dim3(128), 0, st, a); break;
        case 3: hipLaunchKernelGGL(k_render<3>, grid, dim3(128), 0, st, a); break;
        default: hipLaunchKernelGGL(k_render<4>, grid, dim3(128), 0, st, a); break;
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(WX_ERR_DEVICE, std::string("wx_render_camera: launch: ") + hipGetErrorString(e));
    return WX_OK;
}
