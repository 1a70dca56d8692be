// C ABI, host part: errors, the VDB345 tree builder and the camera basis.
// Device-side entry points live in trace.hip.
#include <cmath>
#include <cstring>
#include <exception>
#include <new>

#include "wx_internal.hpp"

namespace wx {

static thread_local std::string g_last_error;

int set_error(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

int ok() { return WX_OK; }

}  // namespace wx

using wx::set_error;

#define WX_GUARD_BEGIN try {
#define WX_GUARD_END                                                          \
    }                                                                         \
    catch (const std::bad_alloc &) {                                          \
        return set_error(WX_ERR_BUDGET, "host allocation failed");            \
    }                                                                         \
    catch (const std::exception &e) {                                         \
        return set_error(WX_ERR_INVALID, std::string("exception: ") + e.what()); \
    }                                                                         \
    catch (...) {                                                             \
        return set_error(WX_ERR_INVALID, "unknown exception");                \
    }

extern "C" {

const char *wx_last_error(void) { return wx::g_last_error.c_str(); }

int wx_version(void) { return 1; }

int wx_vdb_create(wx_vdb **out) {
    WX_GUARD_BEGIN
    if (!out) return set_error(WX_ERR_INVALID, "wx_vdb_create: out is NULL");
    *out = new wx_vdb();
    return WX_OK;
    WX_GUARD_END
}

int wx_vdb_destroy(wx_vdb *v) {
    delete v;
    return WX_OK;
}

int wx_vdb_set_voxel(wx_vdb *v, int32_t x, int32_t y, int32_t z, uint32_t value_bits) {
    WX_GUARD_BEGIN
    if (!v) return set_error(WX_ERR_INVALID, "wx_vdb_set_voxel: NULL tree");
    std::string err;
    if (!v->tree.set_voxel(x, y, z, value_bits, true, err))
        return set_error(v->tree.finalized() ? WX_ERR_STATE : WX_ERR_INVALID, err);
    return WX_OK;
    WX_GUARD_END
}

int wx_vdb_set_voxels(wx_vdb *v, const int32_t *xyz, size_t n) {
    WX_GUARD_BEGIN
    if (!v || (!xyz && n)) return set_error(WX_ERR_INVALID, "wx_vdb_set_voxels: NULL argument");
    std::string err;
    for (size_t i = 0; i < n; ++i)
        if (!v->tree.set_voxel(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], 0, false, err))
            return set_error(v->tree.finalized() ? WX_ERR_STATE : WX_ERR_INVALID, err);
    return WX_OK;
    WX_GUARD_END
}

int wx_vdb_add_spheres(wx_vdb *v, const double *spheres, size_t n, const int32_t lo[3], const int32_t hi[3],
                       int threads) {
    WX_GUARD_BEGIN
    if (!v || (!spheres && n) || !lo || !hi) return set_error(WX_ERR_INVALID, "wx_vdb_add_spheres: NULL argument");
    for (size_t i = 0; i < n; ++i)
        if (!(spheres[4 * i + 3] >= 0.0) || !std::isfinite(spheres[4 * i]) || !std::isfinite(spheres[4 * i + 1]) ||
            !std::isfinite(spheres[4 * i + 2]) || !std::isfinite(spheres[4 * i + 3]))
            return set_error(WX_ERR_INVALID, "wx_vdb_add_spheres: non-finite centre or negative radius");
    std::string err;
    if (!v->tree.add_spheres(spheres, n, lo, hi, threads < 1 ? 1 : threads, err))
        return set_error(v->tree.finalized() ? WX_ERR_STATE : WX_ERR_INVALID, err);
    return WX_OK;
    WX_GUARD_END
}

int wx_vdb_add_noise(wx_vdb *v, uint64_t seed, const int32_t lo[3], const int32_t hi[3], float base_freq,
                     float occupancy, int threads, float *threshold) {
    WX_GUARD_BEGIN
    if (!v || !lo || !hi) return set_error(WX_ERR_INVALID, "wx_vdb_add_noise: NULL argument");
    std::string err;
    if (!v->tree.add_noise(seed, lo, hi, base_freq, occupancy, threads < 1 ? 1 : threads, threshold, err))
        return set_error(v->tree.finalized() ? WX_ERR_STATE : WX_ERR_INVALID, err);
    return WX_OK;
    WX_GUARD_END
}

int wx_vdb_add_leaf_masks(wx_vdb *v, const int32_t *keys, const uint64_t *masks, size_t n) {
    WX_GUARD_BEGIN
    if (!v || ((!keys || !masks) && n)) return set_error(WX_ERR_INVALID, "wx_vdb_add_leaf_masks: NULL argument");
    std::string err;
    if (!v->tree.add_leaf_masks(keys, masks, n, err))
        return set_error(v->tree.finalized() ? WX_ERR_STATE : WX_ERR_INVALID, err);
    return WX_OK;
    WX_GUARD_END
}

int wx_vdb_add_nodes(wx_vdb *v, int level, const int32_t *keys, const uint64_t *masks, size_t n) {
    WX_GUARD_BEGIN
    if (!v || ((!keys || !masks) && n)) return set_error(WX_ERR_INVALID, "wx_vdb_add_nodes: NULL argument");
    if (level != 3 && level != 4 && level != 5) return set_error(WX_ERR_INVALID, "wx_vdb_add_nodes: level 3, 4 or 5");
    std::string err;
    bool ok = true;
    if (level == 3) {
        ok = v->tree.add_leaf_masks(keys, masks, n, err, true);
    } else {
        const size_t words = level == 5 ? 512 : 64;
        for (size_t i = 0; i < n && ok; ++i) ok = v->tree.add_internal(level, keys + 3 * i, masks + words * i, err);
    }
    if (!ok) return set_error(v->tree.finalized() ? WX_ERR_STATE : WX_ERR_INVALID, err);
    return WX_OK;
    WX_GUARD_END
}

int wx_vdb_finalize(wx_vdb *v, int threads) {
    WX_GUARD_BEGIN
    if (!v) return set_error(WX_ERR_INVALID, "wx_vdb_finalize: NULL tree");
    std::string err;
    if (!v->tree.finalize(threads < 1 ? 1 : threads, err)) return set_error(WX_ERR_UNSUPPORTED, err);
    return WX_OK;
    WX_GUARD_END
}

int wx_vdb_get_voxel(const wx_vdb *v, int32_t x, int32_t y, int32_t z, int *kind, uint64_t *value, int *level) {
    if (!v || !kind || !value || !level) return set_error(WX_ERR_INVALID, "wx_vdb_get_voxel: NULL argument");
    if (!v->tree.finalized()) return set_error(WX_ERR_STATE, "wx_vdb_get_voxel: tree not finalized");
    std::string err;
    if (!v->tree.host_words(err)) return set_error(WX_ERR_DEVICE, "wx_vdb_get_voxel: " + err);
    *kind = v->tree.get_voxel(x, y, z, value, level);
    return WX_OK;
}

int wx_vdb_count_nodes(const wx_vdb *v, uint64_t out[3]) {
    if (!v || !out) return set_error(WX_ERR_INVALID, "wx_vdb_count_nodes: NULL argument");
    if (!v->tree.finalized()) return set_error(WX_ERR_STATE, "wx_vdb_count_nodes: tree not finalized");
    v->tree.count_nodes(out);
    return WX_OK;
}

int wx_vdb_active_voxels(const wx_vdb *v, uint64_t *out) {
    if (!v || !out) return set_error(WX_ERR_INVALID, "wx_vdb_active_voxels: NULL argument");
    *out = v->tree.active_voxels();
    return WX_OK;
}

int wx_vdb_ref_masks(const wx_vdb *v, uint32_t *kids5, uint32_t *vals5, uint32_t *kids4, uint32_t *vals4,
                     uint32_t *vals3) {
    if (!v || !kids5 || !vals5 || !kids4 || !vals4 || !vals3)
        return set_error(WX_ERR_INVALID, "wx_vdb_ref_masks: NULL argument");
    if (!v->tree.finalized()) return set_error(WX_ERR_STATE, "wx_vdb_ref_masks: tree not finalized");
    v->tree.ref_masks(kids5, vals5, kids4, vals4, vals3);
    return WX_OK;
}

int wx_vdb_ref_atlas(const wx_vdb *v, uint32_t *atlas5, uint32_t *atlas4, uint32_t *atlas3) {
    if (!v || !atlas5 || !atlas4 || !atlas3) return set_error(WX_ERR_INVALID, "wx_vdb_ref_atlas: NULL argument");
    if (!v->tree.finalized()) return set_error(WX_ERR_STATE, "wx_vdb_ref_atlas: tree not finalized");
    std::string err;
    if (!v->tree.host_words(err)) return set_error(WX_ERR_DEVICE, "wx_vdb_ref_atlas: " + err);
    v->tree.ref_atlas(atlas5, atlas4, atlas3);
    return WX_OK;
}

int wx_vdb_leaf_origins(const wx_vdb *v, int32_t *out) {
    if (!v || !out) return set_error(WX_ERR_INVALID, "wx_vdb_leaf_origins: NULL argument");
    if (!v->tree.finalized()) return set_error(WX_ERR_STATE, "wx_vdb_leaf_origins: tree not finalized");
    const auto &k = v->tree.leaf_keys();
    for (size_t i = 0; i < k.size(); ++i) out[i] = k[i] * 8;
    return WX_OK;
}

int wx_vdb_layout(const wx_vdb *v, wx_grid_desc *out) {
    if (!v || !out) return set_error(WX_ERR_INVALID, "wx_vdb_layout: NULL argument");
    if (!v->tree.finalized()) return set_error(WX_ERR_STATE, "wx_vdb_layout: tree not finalized");
    const wx::HostVdb &t = v->tree;
    std::string err;
    if (!t.host_words(err)) return set_error(WX_ERR_DEVICE, "wx_vdb_layout: " + err);
    out->n5 = t.n5();
    out->n4 = t.n4();
    out->n3 = t.n3();
    out->bound = 4096.0f;
    out->root = t.root_table().data();
    out->n5_records = reinterpret_cast<const uint32_t *>(t.n5_records().data());
    out->n4_records = reinterpret_cast<const uint32_t *>(t.n4_records().data());
    out->leaves = t.leaf_slots();
    return WX_OK;
}

// ComputeState::build (compute_state.rs:87-131) over look_at_rh (camera.rs:31-35):
// f = normalize(target - eye), s = normalize(f x up), u = s x f; the camera-to-world
// columns are (s, u, -f); wp = (-W/2) s + (H/2) u - w (H/2) / tan(fovy/2), mv = -u.
int wx_camera_basis(const float eye[3], const float target[3], const float up[3], float aspect, float fovy_deg,
                    float width, wx_camera *out) {
    if (!eye || !target || !up || !out) return set_error(WX_ERR_INVALID, "wx_camera_basis: NULL argument");
    if (!(aspect > 0.0f) || !(width > 0.0f) || !(fovy_deg > 0.0f && fovy_deg < 180.0f))
        return set_error(WX_ERR_INVALID, "wx_camera_basis: aspect, width and fovy must be positive (fovy < 180)");
    float f[3], s[3], u[3];
    for (int c = 0; c < 3; ++c) f[c] = target[c] - eye[c];
    float fl = std::sqrt((f[0] * f[0] + f[1] * f[1]) + f[2] * f[2]);
    if (!(fl > 0.0f)) return set_error(WX_ERR_INVALID, "wx_camera_basis: eye == target");
    float fi = 1.0f / fl;
    for (int c = 0; c < 3; ++c) f[c] = f[c] * fi;
    s[0] = f[1] * up[2] - f[2] * up[1];
    s[1] = f[2] * up[0] - f[0] * up[2];
    s[2] = f[0] * up[1] - f[1] * up[0];
    float sl = std::sqrt((s[0] * s[0] + s[1] * s[1]) + s[2] * s[2]);
    if (!(sl > 0.0f)) return set_error(WX_ERR_INVALID, "wx_camera_basis: up is parallel to the view direction");
    float si = 1.0f / sl;
    for (int c = 0; c < 3; ++c) s[c] = s[c] * si;
    u[0] = s[1] * f[2] - s[2] * f[1];
    u[1] = s[2] * f[0] - s[0] * f[2];
    u[2] = s[0] * f[1] - s[1] * f[0];
    const float height = width / aspect;
    const float half_w = -width / 2.0f, half_h = height / 2.0f;
    const float rad = fovy_deg * (3.14159265358979323846f / 180.0f);
    const float tn = std::tan(rad * 0.5f);
    for (int c = 0; c < 3; ++c) {
        const float w = -f[c];
        out->eye[c] = eye[c];
        out->u[c] = s[c];
        out->mv[c] = -u[c];
        out->wp[c] = (half_w * s[c] + half_h * u[c]) - (w * half_h) / tn;
    }
    out->eye[3] = out->u[3] = out->mv[3] = out->wp[3] = 0.0f;
    return WX_OK;
}

}  // extern "C"
