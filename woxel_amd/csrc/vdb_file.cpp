// OpenVDB (.vdb) file reader: header, grid descriptors, transform and tree
// topology of a VDB345 float/half grid into the host builder.
//
// Follows VdbReader (src/vdb/read.rs): VdbReader::new (:62-120),
// read_grid_descriptors (:169-211), read_metadata (:213-271), read_transform
// (:141-167), read_tree_topology (:273-347), read_internal_node_header
// (:349-372) and the framing of read_compressed / read_compressed_data
// (:374-519). Only the topology is kept: leaf value masks (active voxels),
// internal value masks (active tiles) and the node structure. The compressed
// value blocks are skipped by their byte counts -- the traversal never reads
// voxel values (an active voxel is a hit whatever its value) -- and the data
// section (read.rs:522-575) repeats the leaf masks, so it is not read. Root
// tiles are read and dropped: the reference's traversal never sees them.
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "wx_internal.hpp"

namespace {

constexpr uint64_t kMagic = 0x56444220ull;
constexpr uint32_t kVersionMin = 218;             // boost UUID
constexpr uint32_t kVersionSelective = 220;       // selective compression byte
constexpr uint32_t kVersionNodeMask = 222;        // node-mask compression
constexpr uint32_t kVersionPerGrid = 223;         // per-grid compression
constexpr uint32_t kZip = 1, kActiveMask = 2, kBlosc = 4;

struct Bytes {
    const uint8_t *p = nullptr;
    size_t n = 0, at = 0;
    std::string err;

    bool need(size_t k) {
        if (at + k > n || at + k < at) {
            if (err.empty()) err = "unexpected end of file at byte " + std::to_string(at);
            return false;
        }
        return true;
    }
    template <typename T>
    bool get(T &v) {
        if (!need(sizeof(T))) return false;
        std::memcpy(&v, p + at, sizeof(T));  // the format is little-endian, like the host
        at += sizeof(T);
        return true;
    }
    bool skip(size_t k) {
        if (!need(k)) return false;
        at += k;
        return true;
    }
    bool str(size_t len, std::string &s) {
        if (!need(len)) return false;
        s.assign(reinterpret_cast<const char *>(p + at), len);
        at += len;
        return true;
    }
    bool lstr(std::string &s) {
        uint32_t len = 0;
        return get(len) && str(len, s);
    }
    bool words(uint64_t *w, size_t k) {
        if (!need(8 * k)) return false;
        std::memcpy(w, p + at, 8 * k);
        at += 8 * k;
        return true;
    }
    bool seek(uint64_t pos) {
        if (pos > n) {
            if (err.empty()) err = "seek past the end of file";
            return false;
        }
        at = (size_t)pos;
        return true;
    }
};

struct Meta {
    int64_t voxel_count = -1;
    bool half = false;
};

// read.rs:213-271; keeps the two entries the reader needs.
bool read_metadata(Bytes &b, Meta *m) {
    uint32_t count = 0;
    if (!b.get(count)) return false;
    for (uint32_t i = 0; i < count; ++i) {
        std::string name, type;
        uint32_t len = 0;
        if (!b.lstr(name) || !b.lstr(type) || !b.get(len)) return false;
        if (m && name == "file_voxel_count" && type == "int64" && len == 8) {
            if (!b.get(m->voxel_count)) return false;
        } else if (m && name == "is_saved_as_half_float" && type == "bool" && len == 1) {
            uint8_t v = 0;
            if (!b.get(v)) return false;
            m->half = v == 1;
        } else if (!b.skip(len)) {
            return false;
        }
    }
    return true;
}

size_t popcount_words(const uint64_t *w, size_t k) {
    size_t c = 0;
    for (size_t i = 0; i < k; ++i) c += (size_t)__builtin_popcountll(w[i]);
    return c;
}

// Framing of one compressed value block (read.rs:374-519), skipped.
bool skip_values(Bytes &b, uint32_t version, uint32_t compression, size_t size, const uint64_t *value_mask,
                 size_t mask_words, size_t elem, std::string &err) {
    uint8_t meta = 6;  // NoMaskAndAllVals before node-mask compression
    if (version >= kVersionNodeMask && !b.get(meta)) return false;
    if (meta > 6) {
        err = "invalid node metadata " + std::to_string(meta);
        return false;
    }
    if (meta == 2 || meta == 4) {  // one inactive value
        if (!b.skip(elem)) return false;
    } else if (meta == 5) {  // two inactive values
        if (!b.skip(2 * elem)) return false;
    }
    if (meta == 3 || meta == 4 || meta == 5)  // selection mask
        if (!b.skip(8 * ((size + 63) / 64))) return false;
    size_t count = size;
    if ((compression & kActiveMask) && meta != 6 && version >= kVersionNodeMask)
        count = popcount_words(value_mask, mask_words);
    if (compression & (kBlosc | kZip)) {
        int64_t bytes = 0;
        if (!b.get(bytes)) return false;
        return b.skip((size_t)(bytes <= 0 ? -bytes : bytes));
    }
    return b.skip(count * elem);
}

}  // namespace

extern "C" int wx_vdb_read_file(wx_vdb *v, const char *path, const char *grid_name, wx_vdb_file_info *info) {
    using wx::set_error;
    try {
        if (!v || !path) return set_error(WX_ERR_INVALID, "wx_vdb_read_file: NULL argument");
        if (v->tree.finalized()) return set_error(WX_ERR_STATE, "wx_vdb_read_file: tree already finalized");
        std::vector<uint8_t> file;
        {
            FILE *f = std::fopen(path, "rb");
            if (!f) return set_error(WX_ERR_INVALID, std::string("wx_vdb_read_file: cannot open ") + path);
            std::fseek(f, 0, SEEK_END);
            const long len = std::ftell(f);
            std::fseek(f, 0, SEEK_SET);
            if (len < 0) {
                std::fclose(f);
                return set_error(WX_ERR_INVALID, "wx_vdb_read_file: cannot size the file");
            }
            file.resize((size_t)len);
            const size_t got = len ? std::fread(file.data(), 1, (size_t)len, f) : 0;
            std::fclose(f);
            if (got != (size_t)len) return set_error(WX_ERR_INVALID, "wx_vdb_read_file: short read");
        }
        Bytes b;
        b.p = file.data();
        b.n = file.size();
        std::string err;
        auto fail = [&](int code, const std::string &what) {
            return set_error(code, "wx_vdb_read_file: " + (what.empty() ? b.err : what));
        };

        // ---- archive header (read.rs:62-120)
        uint64_t magic = 0;
        uint32_t version = 0, lib_major = 0, lib_minor = 0;
        uint8_t has_offsets = 0;
        if (!b.get(magic)) return fail(WX_ERR_INVALID, "");
        if (magic != kMagic) return fail(WX_ERR_INVALID, "not a .vdb file (magic mismatch)");
        if (!b.get(version) || !b.get(lib_major) || !b.get(lib_minor) || !b.get(has_offsets))
            return fail(WX_ERR_INVALID, "");
        if (version < kVersionMin) return fail(WX_ERR_UNSUPPORTED, "file version " + std::to_string(version) + " < 218");
        uint32_t compression = version < kVersionPerGrid ? (kZip | kActiveMask) : (kBlosc | kActiveMask);
        if (version >= kVersionSelective && version < kVersionNodeMask) {
            uint8_t c = 0;
            if (!b.get(c)) return fail(WX_ERR_INVALID, "");
            compression = c == 1 ? kZip : 0u;
        }
        if (!b.skip(36) || !read_metadata(b, nullptr)) return fail(WX_ERR_INVALID, "");  // uuid, file metadata
        uint32_t ngrids = 0;
        if (!b.get(ngrids)) return fail(WX_ERR_INVALID, "");
        if (!has_offsets) return fail(WX_ERR_UNSUPPORTED, "files without grid offsets");

        // ---- grid descriptors (read.rs:169-211): pick the named grid (or the first)
        bool found = false;
        std::string gname;
        uint64_t grid_pos = 0;
        uint32_t gcomp = compression;
        Meta gmeta;
        for (uint32_t i = 0; i < ngrids; ++i) {
            std::string name, type, parent;
            uint64_t gpos = 0, bpos = 0, epos = 0;
            if (!b.lstr(name) || !b.lstr(type) || !b.lstr(parent) || !b.get(gpos) || !b.get(bpos) || !b.get(epos))
                return fail(WX_ERR_INVALID, "");
            uint32_t c = compression;
            if (version >= kVersionNodeMask && !b.get(c)) return fail(WX_ERR_INVALID, "");
            Meta m;
            if (!read_metadata(b, &m)) return fail(WX_ERR_INVALID, "");
            if (!found && (!grid_name || name == grid_name)) {
                found = true;
                gname = name;
                grid_pos = gpos;
                gcomp = c;
                gmeta = m;
            }
            if (!b.seek(epos)) return fail(WX_ERR_INVALID, "");
        }
        if (!found)
            return fail(WX_ERR_INVALID, std::string("no grid named ") + (grid_name ? grid_name : "(any)"));

        // ---- the grid (read.rs:122-139): compression, metadata, transform
        if (!b.seek(grid_pos)) return fail(WX_ERR_INVALID, "");
        if (version >= kVersionNodeMask) {
            uint32_t c = 0;
            if (!b.get(c)) return fail(WX_ERR_INVALID, "");
        }
        if (!read_metadata(b, nullptr)) return fail(WX_ERR_INVALID, "");
        std::string tname;
        if (!b.lstr(tname)) return fail(WX_ERR_INVALID, "");
        double tr[6][3] = {};
        int nvec = 0, vsize_at = 0;
        if (tname == "UniformScaleMap") nvec = 5, vsize_at = 1;
        else if (tname == "UniformScaleTranslateMap" || tname == "ScaleTranslateMap") nvec = 6, vsize_at = 2;
        else return fail(WX_ERR_UNSUPPORTED, "transform " + tname);
        for (int k = 0; k < nvec; ++k)
            for (int a = 0; a < 3; ++a)
                if (!b.get(tr[k][a])) return fail(WX_ERR_INVALID, "");

        // ---- topology (read.rs:273-347)
        uint32_t buffers = 0, background = 0, ntiles = 0, nn5 = 0;
        if (!b.get(buffers) || !b.get(background) || !b.get(ntiles) || !b.get(nn5)) return fail(WX_ERR_INVALID, "");
        if (buffers != 1) return fail(WX_ERR_UNSUPPORTED, "multi-buffer trees");
        for (uint32_t i = 0; i < ntiles; ++i)
            if (!b.skip(12 + 4 + 1)) return fail(WX_ERR_INVALID, "");  // origin, value, active
        const size_t elem = gmeta.half ? 2 : 4;
        std::vector<uint64_t> c5(512), v5(512), c4(64), v4(64), leaf_masks;
        std::vector<int32_t> leaf_keys;
        uint64_t n4count = 0, active_tiles = 0;
        for (uint32_t i = 0; i < nn5; ++i) {
            int32_t o[3];
            if (!b.get(o[0]) || !b.get(o[1]) || !b.get(o[2])) return fail(WX_ERR_INVALID, "");
            const int32_t k5[3] = {o[0] >> 12, o[1] >> 12, o[2] >> 12};  // root_key_from_coords
            if (!b.words(c5.data(), 512) || !b.words(v5.data(), 512)) return fail(WX_ERR_INVALID, "");
            const size_t size5 = version >= kVersionNodeMask ? 32768 : 32768 - popcount_words(c5.data(), 512);
            if (!skip_values(b, version, gcomp, size5, v5.data(), 512, elem, err))
                return fail(WX_ERR_INVALID, err);
            if (!v->tree.add_internal(5, k5, v5.data(), err)) return fail(WX_ERR_INVALID, err);
            active_tiles += popcount_words(v5.data(), 512);
            for (uint32_t w = 0; w < 512; ++w)
                for (uint64_t bits = c5[w]; bits; bits &= bits - 1) {
                    const uint32_t off5 = w * 64 + (uint32_t)__builtin_ctzll(bits);
                    if (!b.words(c4.data(), 64) || !b.words(v4.data(), 64)) return fail(WX_ERR_INVALID, "");
                    const size_t size4 = version >= kVersionNodeMask ? 4096 : 4096 - popcount_words(c4.data(), 64);
                    if (!skip_values(b, version, gcomp, size4, v4.data(), 64, elem, err))
                        return fail(WX_ERR_INVALID, err);
                    const int32_t k4[3] = {k5[0] * 32 + (int32_t)(off5 >> 10), k5[1] * 32 + (int32_t)((off5 >> 5) & 31),
                                           k5[2] * 32 + (int32_t)(off5 & 31)};
                    if (!v->tree.add_internal(4, k4, v4.data(), err)) return fail(WX_ERR_INVALID, err);
                    ++n4count;
                    active_tiles += popcount_words(v4.data(), 64);
                    for (uint32_t w4 = 0; w4 < 64; ++w4)
                        for (uint64_t bits4 = c4[w4]; bits4; bits4 &= bits4 - 1) {
                            const uint32_t off4 = w4 * 64 + (uint32_t)__builtin_ctzll(bits4);
                            const size_t at = leaf_masks.size();
                            leaf_masks.resize(at + 8);
                            if (!b.words(&leaf_masks[at], 8)) return fail(WX_ERR_INVALID, "");
                            leaf_keys.push_back(k4[0] * 16 + (int32_t)(off4 >> 8));
                            leaf_keys.push_back(k4[1] * 16 + (int32_t)((off4 >> 4) & 15));
                            leaf_keys.push_back(k4[2] * 16 + (int32_t)(off4 & 15));
                        }
                }
        }
        const size_t nleaves = leaf_keys.size() / 3;
        if (!v->tree.add_leaf_masks(leaf_keys.data(), leaf_masks.data(), nleaves, err, true))
            return fail(WX_ERR_INVALID, err);
        if (info) {
            std::memset(info, 0, sizeof *info);
            info->file_version = version;
            info->library_major = lib_major;
            info->library_minor = lib_minor;
            info->compression = gcomp;
            info->half_float = gmeta.half ? 1u : 0u;
            info->n5 = nn5;
            info->n4 = (uint32_t)n4count;
            info->n3 = (uint32_t)nleaves;
            info->root_tiles = ntiles;
            info->active_tiles = (uint32_t)active_tiles;
            info->file_voxel_count = gmeta.voxel_count;
            info->active_voxels = popcount_words(leaf_masks.data(), leaf_masks.size());
            for (int a = 0; a < 3; ++a) info->voxel_size[a] = tr[vsize_at][a];
            std::snprintf(info->grid_name, sizeof info->grid_name, "%s", gname.c_str());
        }
        return WX_OK;
    } catch (const std::bad_alloc &) {
        return set_error(WX_ERR_BUDGET, "wx_vdb_read_file: host allocation failed");
    } catch (const std::exception &e) {
        return set_error(WX_ERR_INVALID, std::string("wx_vdb_read_file: ") + e.what());
    }
}
