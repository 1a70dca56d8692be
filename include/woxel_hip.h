/*
 * woxel_hip.h — C ABI of the MI355X-native HDDA + grid-SDF traversal over VDB345.
 *
 * This is the drop-in boundary for woxel's traversal path. The reference exposes
 * the path as a wgpu compute pipeline (src/render/pipelines/compute.rs:6-22) bound
 * to a camera uniform, three u32 atlas textures and six mask/origin storage buffers
 * (src/shaders/raycast.comp.wgsl:22-58, src/render/wgpu_context.rs:101-159, 243-282).
 * A Rust host in src/render would bind these functions in an `extern "C"` block and
 * link libwoxel_hip.so from build.rs.
 *
 * Conventions
 *  - Every function returns an int status: WX_OK (0) or a WX_ERR_* code; the
 *    message of the last failure on the calling thread is wx_last_error().
 *    Nothing unwinds across the ABI.
 *  - Host buffers belong to the caller. Grids (wx_grid) own their device memory.
 *    Ray and hit buffers passed to wx_trace_* are DEVICE pointers owned by the
 *    caller. Calls that take a stream are ordered on that stream (NULL = the
 *    default stream) and return before the work finishes.
 *  - A wx_grid is immutable once created: concurrent traces on one grid are safe.
 */
#ifndef WOXEL_HIP_H
#define WOXEL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WX_OK 0
#define WX_ERR_INVALID 1     /* bad argument */
#define WX_ERR_DEVICE 2      /* HIP runtime error / no device */
#define WX_ERR_UNSUPPORTED 3 /* grid the device layout cannot represent */
#define WX_ERR_STATE 4       /* call not valid in the object's current state */
#define WX_ERR_BUDGET 5      /* device-memory budget exhausted */
#define WX_ERR_BUSY 6        /* the server's request queue is full (back-pressure): retry later */
#define WX_ERR_NOT_FOUND 7   /* layout cache: no file at the path, or one written under another key */

/* Hit states (HDDAout.state, raycast.comp.wgsl:128-142). */
#define WX_HIT 0
#define WX_OOB 1
#define WX_MAXSTEPS 2

/* Hit state of a shadow record whose primary ray did not hit (no shadow ray cast). */
#define WX_NO_RAY 3

/* Trace flags. */
#define WX_TRACE_CLIP 1u   /* entry clip: origins outside [-bound, bound]^3 first move to the box */
#define WX_TRACE_SHADOW 2u /* batch requests: one shadow ray per primary hit toward `light` */
#define WX_TRACE_SIMPLE 4u /* one lane per ray, one launch shape (no persistent work queue); such a
                              wx_trace_camera / wx_trace_rays call may be captured in a HIP graph */

/* Request kinds of a batch. */
#define WX_REQ_CAMERA 0u
#define WX_REQ_RAYS 1u

typedef void *wx_stream; /* a hipStream_t */

/* One ray's result, 32 bytes (the framebuffer element). */
typedef struct wx_hit {
    int32_t voxel[3];  /* i32(floor(p)) at return: the leaf voxel coordinate */
    float t;           /* ((p.x-o.x)*d.x + (p.y-o.y)*d.y) + (p.z-o.z)*d.z */
    uint16_t steps[4]; /* lookups that resolved at tree depth 0 (root miss) / 1 / 2 / 3 */
    uint8_t state;     /* WX_HIT / WX_OOB / WX_MAXSTEPS */
    int8_t normal[3];  /* -sign(d) on the axes of the last step, 0 elsewhere */
    uint32_t lookups;  /* node-5 | node-4 << 10 | node-3 << 20 lookup counts */
} wx_hit;

/* Camera basis words: ComputeState bytes 128..191 (compute_state.rs:9-29):
 * eye, u, mv (= -v), wp, each an f32x4 with w unused. */
typedef struct wx_camera {
    float eye[4];
    float u[4];
    float mv[4];
    float wp[4];
} wx_camera;

/* An explicit ray: origin and direction, 24 bytes. */
typedef struct wx_ray {
    float o[3];
    float d[3];
} wx_ray;

/* One request of a batched launch. */
typedef struct wx_request {
    uint32_t kind;          /* WX_REQ_CAMERA or WX_REQ_RAYS */
    uint32_t grid;          /* index into the batch's grid array */
    uint32_t width, height; /* CAMERA: image size; RAYS: width = ray count, height = 1 */
    uint32_t flags;         /* WX_TRACE_CLIP | WX_TRACE_SHADOW */
    float pixel_offset;     /* CAMERA: 0.001 in the reference */
    wx_camera camera;       /* CAMERA */
    const wx_ray *rays;     /* RAYS: device buffer of width rays */
    float light[4];         /* SHADOW: point light (xyz) */
    wx_hit *out;            /* device, width*height primary records */
    wx_hit *shadow_out;     /* device, width*height shadow records (SHADOW), else NULL */
} wx_request;

/* Launch shape of a traversal (see wx_autotune). */
typedef struct wx_tune {
    uint32_t persistent;   /* 0: one lane per ray, one workgroup per 256 ray slots;
                              1: persistent waves refilled from a work queue (fields below);
                              2: compact -- a lone primary camera request in 16x16 workgroups that
                                 move the rays of their emptiest wave into idle lanes of the others
                                 every `refill` iterations (other batches use shape 0);
                              3: LDS -- a lone camera request (shadow rays too) on a grid with one
                                 node-5: 32x32-pixel workgroups with the node-5 records staged in
                                 LDS (other batches use shape 0) */
    uint32_t refill;       /* 1: refill a wave's idle lanes once at least this many are idle (1..64);
                              2: iterations between workgroup censuses */
    uint32_t chunk;        /* work slots a wave takes from the queue at once (multiple of 64) */
    uint32_t waves_per_cu; /* resident waves per CU launched (1..32) */
    uint32_t xcd_split;    /* shape 1: nonzero = one queue region per XCD, stealing when drained;
                              0 = one region. Shape 0 camera requests: 2 = each XCD takes one
                              contiguous band of the image's 16x8 tiles, else the hardware's
                              round-robin of tiles over the XCDs */
} wx_tune;

typedef struct wx_vdb wx_vdb;   /* host VDB345 tree + SDF + device layout builder */
typedef struct wx_grid wx_grid; /* device-resident grid */

/* Flat device layout of one grid, as produced by wx_vdb_layout (host pointers).
 *   root       n5 x {key.x, key.y, key.z, index}: node-5 keys (origin >> 12), sorted
 *   n5_records n5 x 4096 x 16 B: 8 slots each = {kids8 | walk8 << 8 | vals8 << 16, first child,
 *              sdf bytes[8]}; walk = kids & ~vals, an active tile's byte is 0
 *   n4_records n4 x 256 x 16 B: 16 slots each = {kids16 | walk16 << 16, first child, sdf nibbles[16]};
 *              an active tile's nibble is 0 (inactive tiles hold 1..15)
 *   leaves     n3 x 64 words: 8 nibbles per word; 0 = active voxel, 1..15 = SDF distance
 * Child indices are DFS positions, as in the reference atlases (vdb345.rs:163-264).
 * Limits: n5 <= 65536 and n4 <= 2^20 (4 GiB per node pool); n3 is bounded by memory. */
typedef struct wx_grid_desc {
    uint32_t n5, n4, n3;
    float bound; /* OOB half-extent, in (0, 2^100]; the reference hard-codes 4096 (raycast.comp.wgsl:101) */
    const int32_t *root;
    const uint32_t *n5_records;
    const uint32_t *n4_records;
    const uint32_t *leaves;
} wx_grid_desc;

/* ---------------- errors / device ---------------- */

/* Message of the last failed call on this thread ("" if none). */
/* Message of the last failure on this thread (the reference panics instead, e.g.
 * compute_state.rs:97-100, wgpu_context.rs:101-103). */
const char *wx_last_error(void);
/* ABI version, bumped on incompatible changes. */
int wx_version(void);
/* Number of visible HIP devices (0 on a host without a GPU). */
int wx_device_count(int *count);

/* ---------------- host tree (replaces VDB345<T>, src/vdb/vdb345.rs) ---------------- */

/* VDB::new (data_structure.rs:313-330): an empty tree; destroy frees it. */
int wx_vdb_create(wx_vdb **out);
int wx_vdb_destroy(wx_vdb *v);
/* VDB345::set_voxel (vdb345.rs:26-66); coordinates in [-2^23, 2^23). */
int wx_vdb_set_voxel(wx_vdb *v, int32_t x, int32_t y, int32_t z, uint32_t value_bits);
/* Bulk set_voxel (vdb345.rs:26-66) of n points (xyz: n x 3); values are not stored. */
int wx_vdb_set_voxels(wx_vdb *v, const int32_t *xyz, size_t n);
/* Union of spheres (n x {cx, cy, cz, r}, f64), clipped to lo <= voxel < hi: the
 * sphere-set scenes, equal to set_voxel (vdb345.rs:26-66) of every voxel inside. No
 * counterpart in the reference (its scenes come from .vdb files). */
int wx_vdb_add_spheres(wx_vdb *v, const double *spheres, size_t n, const int32_t lo[3], const int32_t hi[3],
                       int threads);
/* Isosurface of seeded 3-octave value noise (frequency base_freq per voxel, doubling
 * per octave) inside lo <= voxel < hi: voxels where the noise exceeds the quantile
 * threshold giving about `occupancy` of the box are set; *threshold (may be NULL)
 * receives it. A synthetic-scene generator with no counterpart in the reference. */
int wx_vdb_add_noise(wx_vdb *v, uint64_t seed, const int32_t lo[3], const int32_t hi[3], float base_freq,
                     float occupancy, int threads, float *threshold);
/* The same isosurface with the voxels decided on the device (one wave per candidate
 * leaf, 2^24 candidates per launch) and appended to the host tree: identical leaves
 * to wx_vdb_add_noise, for 4096^3 domains that the host would take minutes over.
 * Synchronous on `stream`. Like wx_vdb_add_noise a synthetic-scene generator: the
 * reference builds its grids with set_voxel (vdb345.rs:26-66). */
int wx_vdb_add_noise_device(wx_vdb *v, uint64_t seed, const int32_t lo[3], const int32_t hi[3], float base_freq,
                            float occupancy, wx_stream stream, float *threshold);
/* Merge raw leaf masks: keys n x 3 leaf coordinates (voxel >> 3), masks n x 8 u64
 * (bit o = voxel offset o = x << 6 | y << 3 | z, data_structure.rs:102-108). */
int wx_vdb_add_leaf_masks(wx_vdb *v, const int32_t *keys, const uint64_t *masks, size_t n);
/* Explicit topology, as a .vdb file's topology section declares it
 * (VdbReader::read_tree_topology, read.rs:273-347): level 5, keys = origin >> 12 and
 * masks = 512 u64 value masks (active tiles) per node; level 4, keys = origin >> 7 and
 * 64 u64 per node; level 3, keys = origin >> 3 and 8 u64 active-voxel masks per leaf,
 * where an all-zero mask still creates the leaf. Nodes exist even without children;
 * parents are created as needed. Active tiles are hits for the traversal and ordinary
 * tiles for compute_sdf, as in the reference. */
int wx_vdb_add_nodes(wx_vdb *v, int level, const int32_t *keys, const uint64_t *masks, size_t n);

/* What wx_vdb_read_file read. */
typedef struct wx_vdb_file_info {
    uint32_t file_version, library_major, library_minor;
    uint32_t compression;     /* the grid's flags: 1 zip, 2 active mask, 4 blosc */
    uint32_t half_float;      /* values stored as 16-bit floats */
    uint32_t n5, n4, n3;      /* nodes of the grid */
    uint32_t root_tiles;      /* root-level tiles (dropped, as the reference's traversal ignores them) */
    uint32_t active_tiles;    /* node-5/4 value-mask bits */
    int64_t file_voxel_count; /* the grid's file_voxel_count metadata, -1 if absent */
    uint64_t active_voxels;   /* active leaf voxels read */
    double voxel_size[3];     /* from the grid's transform */
    char grid_name[64];
} wx_vdb_file_info;

/* Replaces VdbReader::new + read_vdb345_grid (read.rs:62-139, 169-347): reads the
 * topology of grid `grid_name` (NULL = the first grid) of an OpenVDB file into the
 * tree -- leaf value masks, internal value masks and node structure, in the
 * reference's file order. Voxel and tile values are skipped (the traversal does not
 * read them). Follow with wx_vdb_finalize. Errors: WX_ERR_INVALID for a malformed or
 * missing file or grid, WX_ERR_UNSUPPORTED for versions < 218, multi-buffer trees,
 * unknown transforms. */
int wx_vdb_read_file(wx_vdb *v, const char *path, const char *grid_name, wx_vdb_file_info *info);

/* Topology in DFS order + VDB345::compute_sdf (vdb345.rs:290-628) + device layout.
 * The tree is read-only afterwards. */
int wx_vdb_finalize(wx_vdb *v, int threads);
/* Where the time of wx_vdb_finalize_device went. */
typedef struct wx_build_info {
    uint64_t launches;   /* SDF kernel launches (one per level-schedule group and pass) */
    double before_sdf_s; /* topology; node-5/4 neighbour tables and level schedules (host) */
    double sdf_device_s; /* the device part: leaf neighbour table and schedules, sweeps, transfers */
    double total_s;      /* the whole call (the rest: range checks and layout packing, host) */
} wx_build_info;
/* wx_vdb_finalize with both passes of VDB345::compute_sdf (vdb345.rs:290-628)
 * replayed on the GPU: one launch per group of the level schedule, one workgroup per
 * node sweeping its slots in wavefronts; the result is identical to the host replay.
 * The leaf words stay in device memory for wx_vdb_grid_create; host queries that
 * read them (wx_vdb_layout, wx_vdb_get_voxel, wx_vdb_ref_atlas) download them once,
 * on first use. info may be NULL. */
int wx_vdb_finalize_device(wx_vdb *v, int threads, wx_stream stream, wx_build_info *info);
/* VDB345::get_voxel (vdb345.rs:69-106) after finalize: *kind = 0 Offs, 1 Leaf,
 * 2 Innr (level 5 or 4), 3 Root, 4 Bkgr. */
int wx_vdb_get_voxel(const wx_vdb *v, int32_t x, int32_t y, int32_t z, int *kind, uint64_t *value, int *level);
/* VDB345::count_nodes (vdb345.rs:266-287): {N5, N4, N3}. */
int wx_vdb_count_nodes(const wx_vdb *v, uint64_t out[3]);
/* Active leaf voxels (the quantity of file_voxel_count, read.rs:796-806). */
int wx_vdb_active_voxels(const wx_vdb *v, uint64_t *out);
/* VDB345::masks (vdb345.rs:119-161) into caller buffers of n5*1024, n5*1024,
 * n4*128, n4*128, n3*16 u32 words. */
int wx_vdb_ref_masks(const wx_vdb *v, uint32_t *kids5, uint32_t *vals5, uint32_t *kids4, uint32_t *vals4,
                     uint32_t *vals3);
/* VDB345::atlas (vdb345.rs:163-264) slot values, one block per node in index
 * order (n5*32768, n4*4096, n3*512 u32), slots in child-offset order. */
int wx_vdb_ref_atlas(const wx_vdb *v, uint32_t *atlas5, uint32_t *atlas4, uint32_t *atlas3);
/* Origins of the n3 leaves in pool (DFS) order, n3 x 3 int32 (the leaf behind
 * row i of wx_vdb_ref_masks' vals3). */
int wx_vdb_leaf_origins(const wx_vdb *v, int32_t *out);
/* The device layout (the packed form of origins + masks + atlas, vdb345.rs:108-264);
 * pointers stay valid until wx_vdb_destroy. */
int wx_vdb_layout(const wx_vdb *v, wx_grid_desc *out);

/* ---------------- camera (ComputeState::build, compute_state.rs:87-131) ---------------- */

int wx_camera_basis(const float eye[3], const float target[3], const float up[3], float aspect, float fovy_deg,
                    float width, wx_camera *out);

/* ---------------- device grid (replaces the atlas textures and mask buffers,
 *                  src/render/wgpu_context.rs:101-159) ---------------- */

/* Copies the layout to the device: the atlas write_texture calls and mask buffers of
 * wgpu_context.rs:101-159. Destroy replaces the buffer drop on model change
 * (wgpu_context.rs:525-572). */
int wx_grid_create(const wx_grid_desc *desc, wx_grid **out);
/* wx_grid_create over a finalized tree's own layout, with the given bound. After
 * wx_vdb_finalize_device the leaf words are still in device memory (the replay
 * wrote them there) and are copied device to device instead of down to the host
 * and up again; the tree keeps its copy until wx_vdb_destroy. Other trees, or a
 * grid on another device than the replay's, take the host copy. */
int wx_vdb_grid_create(const wx_vdb *v, float bound, wx_grid **out);
int wx_grid_destroy(wx_grid *g);
/* Device bytes held by the grid. */
int wx_grid_bytes(const wx_grid *g, uint64_t *out);

/* ---------------- grid-layout cache ----------------
 * The reference rebuilds its tree and compute_sdf from the scene on every run and has
 * no saved form of the SDF (its .vdb writer, src/vdb/write.rs:10-41, stores topology
 * only). These store a finalized layout (root table, node records with their tile
 * distances, leaf words) under a caller's key -- a string up to 2048 bytes naming what
 * built it, e.g. the scene config and seed -- so a rebuilt grid is the same bytes.
 * One file per layout: a 4 KiB header (format, counts, bound, key, XXH64 digests of
 * every section) and the four pools at 4 KiB-aligned offsets. Writes go to
 * path + ".tmp" and are renamed into place when complete. */
/* Writes a host layout (e.g. wx_vdb_layout's). */
int wx_layout_save(const wx_grid_desc *desc, const char *path, const char *key);
/* Writes a device grid's layout, streamed through pinned staging (no host copy of
 * the pools). */
int wx_grid_save(const wx_grid *g, const char *path, const char *key);
/* Checks the header of a layout file: WX_OK with counts = {n5, n4, n3};
 * WX_ERR_NOT_FOUND for no file or another key; WX_ERR_INVALID for another format, a
 * damaged header or a truncated file. Does not read the pools. */
int wx_layout_probe(const char *path, const char *key, uint32_t counts[3]);
/* Creates a device grid from a layout file written under the same key: the node
 * sections are read, checked against their digests and by wx_grid_create's bounds
 * checks; the leaf words stream to the device through pinned staging, each piece
 * hashed on the way. Errors as wx_layout_probe, and WX_ERR_INVALID (no grid) when
 * any section does not match its digest. */
int wx_grid_load(const char *path, const char *key, wx_grid **out);

/* ---------------- traversal (replaces cp_main + hdda_ray, raycast.comp.wgsl:60-126) -------- */

/* One primary ray per pixel: p = (x, y) + pixel_offset (0.001 in the reference,
 * raycast.comp.wgsl:62), dir = normalize(p.x*u + p.y*mv + wp), origin = eye.
 * out: device buffer of w*h wx_hit, row-major. Unlike the reference dispatch
 * (wgpu_context.rs:281) edge pixels are not dropped when w%8 or h%4 != 0. */
int wx_trace_camera(wx_grid *g, const wx_camera *cam, uint32_t w, uint32_t h, float pixel_offset, uint32_t flags,
                    wx_hit *out, wx_stream stream);
/* hdda_ray (raycast.comp.wgsl:84-126) on explicit rays: rays and out are device
 * buffers of n elements. */
int wx_trace_rays(wx_grid *g, const wx_ray *rays, uint32_t n, uint32_t flags, wx_hit *out, wx_stream stream);

/* ---------------- shading (replaces ray_trace / reflect_ray2 / reflect_ray1 and the
 *                  rgba8 texture store of cp_main, raycast.comp.wgsl:60-68, 144-342) ---------------- */

/* The shading words of ComputeState (compute_state.rs:9-29, 109-117). */
typedef struct wx_shade {
    uint32_t mode;        /* RenderMode (egui_dev.rs:12-18): 0 gray, 1 rgb, 2 ray length, 3 diffuse, 4 glossy */
    uint32_t show_345[3]; /* grid overlay on node-3 / node-4 / node-5 boundaries (0 or 1) */
    float sun_dir[4];     /* direction the sunlight travels; shadow rays go toward -sun_dir */
    float sun_color[4];   /* rgb, intensity */
} wx_shade;

#define WX_RENDER_F32 8u  /* write float4 colours (rgba32f) instead of rgba8 */

/* Colour of each pixel as the reference's compute pass writes it: the primary ray
 * of wx_trace_camera, then per render mode the shadow ray toward the sun (diffuse,
 * glossy) and two reflection bounces with their own shadow rays (glossy), up to six
 * traversals per pixel. out: device buffer of w*h rgba8 words (r | g << 8 | b << 16
 * | 255 << 24, each channel clamped to [0, 1], times 255, rounded to nearest even),
 * or w*h float4 with WX_RENDER_F32. flags: WX_TRACE_CLIP, WX_RENDER_F32. */
int wx_render_camera(wx_grid *g, const wx_camera *cam, uint32_t w, uint32_t h, float pixel_offset,
                     const wx_shade *shade, uint32_t flags, void *out, wx_stream stream);

/* ---------------- batching (many cameras / ray sets / grids in one launch) ----------------
 * No counterpart in the reference, which dispatches one camera per frame
 * (wgpu_context.rs:270-282): these calls batch that dispatch. */

/* The tuning used when none is given (the last wx_autotune result on this device,
 * else built-in defaults). */
int wx_tune_get(wx_tune *out);
int wx_tune_set(const wx_tune *t);
/* All requests in ONE launch. Every request is cut into 64-ray slots (an 8x8
 * pixel tile of a camera, or 64 consecutive rays); waves take slots across
 * requests and grids, and a shadow ray continues on-device in the lane of its
 * primary. The request table is copied to the device asynchronously on
 * `stream`; grids must stay alive until the work completes.
 * HIP graph capture: a call with one request on one grid in shape 0, 2 or 3
 * launches one kernel whose arguments are all passed by value, and may be captured
 * and replayed. Other calls (several requests or grids, shape 1) stage tables and
 * queue counters in per-stream scratch memory that later calls reuse or grow, and
 * must not be captured. */
int wx_trace_batch(wx_grid *const *grids, uint32_t n_grids, const wx_request *reqs, uint32_t n_reqs,
                   const wx_tune *tune, wx_stream stream);
/* wx_trace_batch without any allocation: no counterpart in the reference. For callers
 * that cannot allocate device memory -- code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc. The same kernels and results as
 * wx_trace_batch; the batch tables and queue counters live in one static device array
 * per GPU (140,544 bytes, held until the process ends) instead of per-stream scratch,
 * so at most 64 grids and 1024 requests per call (WX_ERR_INVALID beyond). Calls that
 * overlap on one GPU (other streams, other threads) share that array and race:
 * serialise them. The tables are copied from a per-GPU host staging area with
 * hipMemcpyAsync from pageable memory; the queue counters are cleared by a kernel. */
int wx_trace_batch_static(wx_grid *const *grids, uint32_t n_grids, const wx_request *reqs, uint32_t n_reqs,
                          const wx_tune *tune, wx_stream stream);
/* Times candidate launch shapes on the given batch (synchronously, on `stream`)
 * and stores the fastest as this device's default; *best receives it. */
int wx_autotune(wx_grid *const *grids, uint32_t n_grids, const wx_request *reqs, uint32_t n_reqs,
                wx_stream stream, wx_tune *best, float *best_ms);

/* ---------------- serving (continuous batching over resident grids) ----------------
 * No counterpart in the reference (a single-window renderer): a request loop over
 * the batching calls above. */

typedef struct wx_server wx_server;

#define WX_SERVER_DRYRUN 1u   /* scheduler only: no device work, no outputs (admission/queueing tests) */
#define WX_SERVER_RESIDENT 2u /* admission into a running kernel instead of one launch per batch (below) */

typedef struct wx_server_config {
    uint64_t hbm_budget;         /* device bytes for request outputs (one arena, sub-allocated) */
    uint32_t streams;            /* batches in flight at once, one HIP stream each (1..16) */
    uint32_t max_batch_requests; /* requests per batched launch (1..1024) */
    uint64_t max_batch_rays;     /* primary rays per batched launch (0 = unlimited) */
    uint32_t flags;              /* WX_SERVER_DRYRUN, WX_SERVER_RESIDENT */
    uint32_t max_queued;         /* requests waiting for launch beyond which wx_server_submit returns
                                    WX_ERR_BUSY (0 = no limit) */
    uint32_t wait_limit_us;      /* resident: longest time a wave of the running kernel waits for work
                                    before it ends (0 = 2 s) */
    uint32_t idle_stop_us;       /* resident: the kernel is ended after this long with nothing queued
                                    or in flight (0 = 50 ms) */
} wx_server_config;

typedef struct wx_serve_request {
    uint32_t grid;      /* id from wx_server_add_grid */
    uint32_t width, height;
    uint32_t flags;     /* WX_TRACE_CLIP | WX_TRACE_SHADOW */
    float pixel_offset;
    float light[3];
    wx_camera camera;
} wx_serve_request;

typedef struct wx_serve_result {
    const wx_hit *primary; /* device, width*height records (valid until wx_server_release) */
    const wx_hit *shadow;  /* device, or NULL without WX_TRACE_SHADOW */
    uint32_t width, height;
    double submit_s, launch_s, done_s; /* host monotonic clock */
    uint32_t batch_requests;           /* requests in the launch that served it */
    int status;                        /* WX_OK or the launch's error */
} wx_serve_result;

typedef struct wx_server_stats {
    uint64_t submitted, completed, rejected, cancelled, batches, primary_rays;
    double p50_ms, p99_ms, mean_ms, mean_batch_requests;
    uint64_t budget_in_use, budget_peak;
} wx_server_stats;

/* A serving loop: a dispatcher thread packs queued requests into one batched
 * launch per free stream (requests that arrive while others run join the next
 * launch; at most 64 distinct grids and 2^31 ray slots per launch), admits requests
 * while their outputs fit in the HBM budget, and marks them done when their launch
 * completes. Latency percentiles come from a fixed log-spaced histogram (bins 1 %
 * wide from 1 us to 1000 s). Thread-safe.
 *
 * WX_SERVER_RESIDENT: one kernel stays running while requests arrive (a session) and
 * the dispatcher admits each request into it: it writes the request into a mailbox in
 * pinned host memory and publishes its 8x8 tiles, which the kernel's waves claim as
 * they free up; the kernel flags each request whose outputs are all in memory, and a
 * completion thread marks it done. Every wait in the kernel ends when the host says
 * stop or after wait_limit_us; the host ends a session after idle_stop_us without
 * work, before admitting a grid registered after the session began, and after 16384
 * requests; requests left unfinished by a session that ended are run again in the
 * next one. Sessions take camera requests over at most 64 grids (more: WX_ERR_UNSUPPORTED
 * from wx_server_add_grid); streams and max_batch_* do not apply; batch_requests and
 * the batch statistics count sessions. */
int wx_server_create(const wx_server_config *cfg, wx_server **out);
/* Cancels requests not yet launched, waits for launched ones, frees the arena and streams. */
int wx_server_destroy(wx_server *s);
/* Registers a resident grid (not owned; keep it alive until wx_server_destroy). */
int wx_server_add_grid(wx_server *s, wx_grid *g, uint32_t *id);
/* Queues a request; WX_ERR_BUDGET if its outputs can never fit the budget, WX_ERR_BUSY
 * if max_queued requests are already waiting for launch. */
int wx_server_submit(wx_server *s, const wx_serve_request *r, uint64_t *ticket);
/* Blocks until the request is done (timeout_s < 0: no limit); WX_ERR_STATE on timeout. */
int wx_server_wait(wx_server *s, uint64_t ticket, double timeout_s, wx_serve_result *out);
/* Returns a done request's output buffers to the budget. */
int wx_server_release(wx_server *s, uint64_t ticket);
/* Removes a still-queued request (WX_ERR_STATE once it has been launched). */
int wx_server_cancel(wx_server *s, uint64_t ticket);
int wx_server_get_stats(wx_server *s, wx_server_stats *out);

#ifdef __cplusplus
}
#endif
#endif
