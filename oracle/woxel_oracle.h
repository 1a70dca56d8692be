/*
 * woxel_oracle.h — CPU oracle for the HDDA + grid-SDF traversal over VDB345.
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker, not the product: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it. The
 * product library (woxel_amd/libwoxel_hip.so) never links or calls it.
 *
 * It is a scalar C restatement of the reference algorithm, following the
 * reference data structures (a root map of N5 nodes, InternalNode/LeafNode
 * arrays, the five mask arrays and the three atlases) rather than the
 * product's device layout, so that the product's layout builder and kernels
 * are checked against an independent implementation.
 *
 * Reference files restated (paths relative to the woxel repository):
 *   src/vdb/data_structure.rs:17-92, 264-268   node index math
 *   src/vdb/vdb345.rs:26-106                    set_voxel / get_voxel
 *   src/vdb/vdb345.rs:108-287, 673-694          origins / masks / atlas / count_nodes
 *   src/vdb/vdb345.rs:290-670                   compute_sdf (2-pass, order-dependent)
 *   src/render/gpu_types/mask.rs:96-118         mask buffers + origins
 *   src/shaders/raycast.comp.wgsl:60-126        cp_main ray generation, hdda_ray loop
 *   src/shaders/raycast.comp.wgsl:344-519       VdbLeaf cursor cache and lookups
 *   src/shaders/raycast.comp.wgsl:183-190       shadow-ray spawn
 *   src/shaders/raycast.comp.wgsl:144-342       ray_trace / reflect_ray2 / reflect_ray1 shading
 *   src/vdb/read.rs:303-340                     explicit internal nodes of a .vdb topology
 *   src/render/gpu_types/compute_state.rs:87-131 camera basis (u, mv, wp)
 *
 * Compile with -ffp-contract=off and without fast-math: every FP32 operation
 * is written out in the reference's evaluation order.
 *
 * What pins each part (tests/test_oracle.py, tests/test_layout.py,
 * tests/test_shade_oracle.py, tests/test_vdb_file.py): the reference's index-math
 * and set/get_voxel known-answer tests; compute_sdf against the chessboard distance
 * transform on lone nodes and a literal pure-Python replay on small trees; layout
 * invariants; traversal closed forms (axis scans, the step cap beside a wall, start
 * inside a voxel, empty space) and an f64 voxel walk; shading closed forms on box
 * scenes. Parity unpinned: the camera basis against cgmath 0.18's look_at_rh +
 * Matrix4::invert bits (no Rust toolchain here; the basis is a boundary input that
 * the kernels and this oracle both consume), and NaN sign bits of t for degenerate
 * rays other than those the tests list.
 */
#ifndef WOXEL_ORACLE_H
#define WOXEL_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- reference-style VDB345 tree ---------------- */

typedef struct or_vdb or_vdb;

/* VdbEndpoint kinds (data_structure.rs:338-344). */
enum { OR_EP_OFFS = 0, OR_EP_LEAF = 1, OR_EP_INNR = 2, OR_EP_ROOT = 3, OR_EP_BKGR = 4 };

or_vdb *or_vdb_new(void);
void or_vdb_free(or_vdb *v);
/* VDB345::set_voxel (vdb345.rs:26-66); value_bits is the f32/u32 value payload. */
void or_vdb_set_voxel(or_vdb *v, int32_t x, int32_t y, int32_t z, uint32_t value_bits);
/* VDB345::get_voxel (vdb345.rs:69-106). Returns the VdbEndpoint kind; *value
 * receives the payload (Offs offset, Leaf value bits, Innr/Root tile value) and
 * *level the Innr level tag (5 or 4). */
int or_vdb_get_voxel(const or_vdb *v, int32_t x, int32_t y, int32_t z, uint64_t *value, int *level);
/* VDB345::compute_sdf (vdb345.rs:290-628), literal forward + backward DFS. */
void or_vdb_compute_sdf(or_vdb *v);
/* VDB345::count_nodes (vdb345.rs:266-287): out = {N5, N4, N3}. */
void or_vdb_count_nodes(const or_vdb *v, uint64_t out[3]);
/* Generator: set every integer voxel with (x-cx)^2+(y-cy)^2+(z-cz)^2 <= r^2
 * (f64, left-to-right sum) inside [lo, hi). value_mode 0: value_bits as given;
 * 1: f32 bits of (r - |x - c|), the distance to the sphere surface. */
void or_vdb_add_sphere(or_vdb *v, double cx, double cy, double cz, double r,
                       const int32_t lo[3], const int32_t hi[3], int value_mode,
                       uint32_t value_bits);
/* set_voxel for every set bit of n leaf value masks (origins n x 3, masks n x 8 u64,
 * bit o = voxel offset o), value_bits as the payload. */
void or_vdb_add_leaf_masks(or_vdb *v, const int32_t *origins, const uint64_t *masks, size_t n, uint32_t value_bits);
/* A node as a .vdb file's topology declares it (read.rs:303-340): level 5 / 4 / 3
 * with its value mask (512 / 64 / 8 u64); exists even without children. */
void or_vdb_add_node(or_vdb *v, int level, const int32_t origin[3], const uint64_t *vmask);
/* Read back one leaf's slots (for tests): data[512] after compute_sdf,
 * isval bitmask (LeafData::Value), returns 0 if the leaf does not exist. */
int or_vdb_leaf(const or_vdb *v, int32_t x, int32_t y, int32_t z, uint64_t data[512],
                uint64_t isval[8]);

/* ---------------- reference GPU layout (vdb345.rs:108-264, mask.rs) ---------------- */

typedef struct or_layout {
    uint32_t n5, n4, n3;
    int32_t *origins; /* n5 x 4 (x, y, z, 0): vec3<i32> with 16-byte stride */
    uint32_t *kids5;  /* n5 x 1024 */
    uint32_t *vals5;  /* n5 x 1024 */
    uint32_t *kids4;  /* n4 x 128 */
    uint32_t *vals4;  /* n4 x 128 */
    uint32_t *vals3;  /* n3 x 16 */
    /* Atlas texels, one block per node in node-index order, slots in child
     * offset order: texel (32*o + c) of the cubic atlas for node idx with
     * o = atlas_origin_from_idx(idx) maps to atlas5[idx*32768 + child_to_offset(c)]. */
    uint32_t *atlas5; /* n5 x 32768 */
    uint32_t *atlas4; /* n4 x 4096 */
    uint32_t *atlas3; /* n3 x 512 */
} or_layout;

or_layout *or_layout_build(const or_vdb *v);
/* A layout from caller arrays in the same format (copied), e.g. a reference-format
 * export of another builder; lets the oracle traverse grids it did not build. */
or_layout *or_layout_from_arrays(uint32_t n5, uint32_t n4, uint32_t n3, const int32_t *origins,
                                 const uint32_t *kids5, const uint32_t *vals5, const uint32_t *kids4,
                                 const uint32_t *vals4, const uint32_t *vals3, const uint32_t *atlas5,
                                 const uint32_t *atlas4, const uint32_t *atlas3);
void or_layout_free(or_layout *l);

/* ---------------- traversal ---------------- */

enum { OR_HIT = 0, OR_OOB = 1, OR_MAXSTEPS = 2 };

/* Hit record, 32 bytes; identical layout to the product's wx_hit. */
typedef struct or_hit {
    int32_t voxel[3];   /* i32(floor(p)) at return */
    float t;            /* ((p.x-o.x)d.x + (p.y-o.y)d.y) + (p.z-o.z)d.z */
    uint16_t steps[4];  /* lookups resolved at num_parents 0/1/2/3 */
    uint8_t state;      /* OR_HIT / OR_OOB / OR_MAXSTEPS */
    int8_t normal[3];   /* -step * mask */
    uint32_t lookups;   /* node5 | node4 << 10 | node3 << 20 lookup invocations */
} or_hit;

/* Diagnostics the oracle alone produces. */
typedef struct or_diag {
    float p[3];     /* final p */
    uint32_t flags; /* OR_F_* */
} or_diag;

enum {
    OR_F_TIE = 1u,       /* some iteration had >1 axis at the minimum tMax (exact-tie grazing) */
    OR_F_NONFINITE = 2u, /* a non-finite tMax or p occurred */
    OR_F_ZERODIR = 4u,   /* a direction component is +-0 */
    OR_F_CLIPPED = 8u,   /* entry clip moved the origin */
    OR_F_CLIPMISS = 16u, /* entry clip: the ray misses the bound box */
};

/* Trace options. bound: OOB half-extent (reference: 4096, raycast.comp.wgsl:101).
 * clip: 1 = entry-clip extension (origins outside [-bound, bound]^3 are first
 * advanced to the box), 0 = reference semantics. */
void or_hdda_ray(const or_layout *l, const float src[3], const float dir[3], float bound, int clip,
                 or_hit *hit, or_diag *diag);

/* Shading words of ComputeState (same layout as the product's wx_shade). */
typedef struct or_shade {
    uint32_t mode;        /* 0 gray, 1 rgb, 2 ray length, 3 diffuse, 4 glossy */
    uint32_t show_345[3]; /* grid overlay on node-3 / -4 / -5 boundaries */
    float sun_dir[4];
    float sun_color[4];   /* rgb, intensity */
} or_shade;

/* ray_trace (raycast.comp.wgsl:152-242) of the pixel's primary ray, with the
 * shadow and reflection rays it casts (reflect_ray2/1, :244-335). */
void or_render_pixel(const or_layout *l, const float cam[16], uint32_t px, uint32_t py, float offset, float bound,
                     int clip, const or_shade *s, float rgb[3]);
/* All pixels; rgba (w*h*4 floats) and/or rgba8 (w*h words) may be NULL. */
void or_render_camera(const or_layout *l, const float cam[16], uint32_t w, uint32_t h, float offset, float bound,
                      int clip, const or_shade *s, float *rgba, uint32_t *rgba8, int nthreads);

/* cp_main ray direction (raycast.comp.wgsl:62-63). cam = eye[4], u[4], mv[4], wp[4]. */
void or_camera_ray(const float cam[16], uint32_t px, uint32_t py, float offset, float dir[3]);

/* Camera basis (compute_state.rs:87-131 with camera.rs:31-35): writes eye, u, mv, wp. */
void or_camera_basis(const float eye[3], const float target[3], const float up[3], float aspect,
                     float fovy_deg, float width, float cam[16]);

/* Whole-image and ray-list drivers (row-major output, pthreads). diag may be NULL. */
void or_trace_camera(const or_layout *l, const float cam[16], uint32_t w, uint32_t h, float offset,
                     float bound, int clip, or_hit *out, or_diag *diag, int nthreads);
/* rays: n x 6 floats (ox, oy, oz, dx, dy, dz). */
void or_trace_rays(const or_layout *l, const float *rays, size_t n, float bound, int clip,
                   or_hit *out, or_diag *diag, int nthreads);

/* Shadow ray from a primary hit (raycast.comp.wgsl:188-190) toward a point light:
 * origin = p - (4e-2 * step) * mask; dir = normalize(light - origin). */
void or_shadow_ray(const float p[3], const float dir[3], const int8_t normal[3], const float light[3],
                   float o_out[3], float d_out[3]);

/* Index math restated for the known-answer tests (data_structure.rs:17-92). */
void or_global_to_node(const int32_t g[3], int total_log2, int32_t out[3]);
uint32_t or_global_to_offset(const int32_t g[3], int log2_d, int total_log2);
void or_offset_to_child(uint32_t off, int log2_d, uint32_t out[3]);
uint32_t or_child_to_offset(const uint32_t c[3], int log2_d);

#ifdef __cplusplus
}
#endif
#endif
