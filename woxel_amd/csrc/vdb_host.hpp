// Host-side VDB345 builder: topology, grid SDF and the device layout.
//
// Mirrors the reference's VDB345<T> (src/vdb/vdb345.rs, src/vdb/data_structure.rs):
// set_voxel/get_voxel semantics, DFS node numbering (sorted root keys, then child
// offsets ascending) and the order-dependent two-pass compute_sdf. The storage is
// MI355X-first: leaves live in one flat pool in DFS order, the SDF replay works on
// per-level arrays with precomputed neighbour-node tables, and the device layout is
// packed as 16-byte node records and 4-bit leaf slots (see pack_layout()).
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

namespace wxnoise {
struct Params;
}

namespace wx {

// std::allocator that leaves new elements of trivial types uninitialised: the large
// build arrays are written in full by their producers, several threads touching the
// pages first instead of one thread zeroing them.
template <typename T>
struct NoInit : std::allocator<T> {
    template <typename U>
    struct rebind {
        using other = NoInit<U>;
    };
    NoInit() = default;
    template <typename U>
    NoInit(const NoInit<U> &) {}
    template <typename U>
    void construct(U *p) noexcept {
        ::new ((void *)p) U;
    }
    template <typename U, typename... A>
    void construct(U *p, A &&...a) {
        ::new ((void *)p) U(std::forward<A>(a)...);
    }
};
template <typename T>
using BigVec = std::vector<T, NoInit<T>>;

// Leaf slot code in the device layout: 0 = active voxel, 1..15 = SDF distance.
// Node-5 records cover 8 slots, node-4 records 16 slots, both 16 bytes.
struct Rec16 {
    uint32_t w[4];
};

// Leaf layout words (n3 x 64 u32, DFS order) that a device replay left in device
// memory. The tree keeps them for the device grid and fetches its host copy from
// them only when a host query needs it.
struct DeviceWords {
    virtual ~DeviceWords() = default;
    virtual const void *data() const = 0;  // device pointer
    virtual int device() const = 0;        // HIP device index holding it
    virtual bool download(uint32_t *dst, size_t words, std::string &err) const = 0;
};

// A buffer in device memory that the tree holds for a device build (the leaves' masks
// a device generator left there).
struct DeviceBuffer {
    virtual ~DeviceBuffer() = default;
    virtual const void *data() const = 0;  // device pointer
    virtual int device() const = 0;        // HIP device index holding it
};

// Distance (SDF) work of one node level, for a replay outside this class (the
// device build): the level's solid slots, its neighbour-node table and the
// level schedules of both passes; `val` is filled by the replay.
struct SdfLevelJob {
    int log = 0;                      // 5, 4 or 3: the node has 2^(3 log) slots
    uint32_t nodes = 0;
    const uint64_t *solid = nullptr;  // nodes x 2^(3 log) / 64 words: child / active-voxel bits
    const void *solid_dev = nullptr;  // leaves: the same words already in device memory, or NULL
    int solid_dev_device = -1;        // the HIP device holding solid_dev
    uint8_t *val = nullptr;           // nodes x 2^(3 log), 255 (unknown) on entry; NULL when words is set
    bool words = false;               // leaves only: the replay produces the packed layout words (64 per
                                      // leaf, 8 4-bit codes each) instead of val, range-checks them and
                                      // leaves them in device memory (kept)
    std::unique_ptr<DeviceWords> kept;
    BigVec<int32_t> nb;               // nodes x 27: index of the node at key + d, or -1
    std::vector<uint32_t> order[2];   // [1] forward, [0] backward: nodes grouped by schedule level
    std::vector<uint32_t> start[2];   // group g = order[start[g] .. start[g + 1])
    // Leaves: when nb is empty, the replay derives the neighbour table and the
    // schedules itself from the node keys (3 per node, DFS order).
    const int32_t *keys = nullptr;
};
// Replays both passes of all three levels (jobs[0] = node-5, [1] = node-4, [2] = leaves).
typedef bool (*SdfRunner)(SdfLevelJob *jobs, void *ctx, std::string &err);

// VdbEndpoint kinds (data_structure.rs:338-344).
enum Endpoint : int { EP_OFFS = 0, EP_LEAF = 1, EP_INNR = 2, EP_ROOT = 3, EP_BKGR = 4 };

class HostVdb {
  public:
    HostVdb();
    ~HostVdb();

    // ---- construction (vdb345.rs:26-66 set_voxel semantics) ----
    // Voxel coordinates must lie in [-2^23, 2^23).
    bool set_voxel(int32_t x, int32_t y, int32_t z, uint32_t value_bits, bool store_value, std::string &err);
    // Union of spheres: voxel (x,y,z) is active iff (x-cx)^2+(y-cy)^2+(z-cz)^2 <= r^2
    // (f64, left-to-right sum) and lo <= (x,y,z) < hi. spheres = n x (cx, cy, cz, r).
    bool add_spheres(const double *spheres, size_t n, const int32_t lo[3], const int32_t hi[3], int threads,
                     std::string &err);
    // Isosurface of seeded 3-octave value noise: voxel (x,y,z) in [lo, hi) is active
    // iff noise(x, y, z) > t, with t chosen so that about `occupancy` of the domain
    // is active (quantile of 65536 seeded samples); *threshold receives t.
    bool add_noise(uint64_t seed, const int32_t lo[3], const int32_t hi[3], float base_freq, float occupancy,
                   int threads, float *threshold, std::string &err);
    // The checked parameters and threshold of that isosurface (for a generator
    // elsewhere, the device one: same voxels).
    bool noise_params(uint64_t seed, const int32_t lo[3], const int32_t hi[3], float base_freq, float occupancy,
                      wxnoise::Params &p, std::string &err) const;
    // Append leaves the caller guarantees are new: keys unique and absent from the
    // tree, masks non-zero. No hashing; the key index is rebuilt on the next lookup.
    bool append_new_leaves(const int32_t *keys, const uint64_t *masks, size_t n, std::string &err);
    // The same for leaves the caller writes itself: grows the pool by n leaves, left
    // uninitialised, and returns where their keys (n x 3) and masks (n x 8 u64) go. The
    // caller fills every one of them before any other call on the tree.
    bool grow_new_leaves(size_t n, int32_t **keys, uint64_t **masks, std::string &err);
    // Takes back the leaves from pool index `base` on (those grow_new_leaves added, when
    // the caller could not fill them).
    void drop_leaves_from(size_t base);
    size_t leaf_count() const { return leaves_.size(); }
    // Merge raw leaf masks: keys = n x 3 leaf coordinates (voxel >> 3), masks = n x 8 u64.
    // keep_empty creates a leaf for an all-zero mask too (a file's topology can hold
    // leaves without active voxels; set_voxel never makes one).
    bool add_leaf_masks(const int32_t *keys, const uint64_t *masks, size_t n, std::string &err,
                        bool keep_empty = false);
    // An internal node as a file's topology declares it (read.rs:303-340): level 5
    // (key = origin >> 12, value mask of 512 u64) or level 4 (key = origin >> 7, 64 u64).
    // The node exists even without children; value-mask bits are active tiles, which
    // compute_sdf treats like any tile and the lookup reports as hits.
    bool add_internal(int level, const int32_t key[3], const uint64_t *value_mask, std::string &err);

    // ---- finalisation: DFS order, child masks, SDF (vdb345.rs:290-628), layout ----
    bool finalize(int threads, std::string &err);
    // The same, with the grid SDF replayed by `run` (bit-identical by contract).
    bool finalize_with(int threads, SdfRunner run, void *ctx, std::string &err);
    bool finalized() const { return finalized_; }

    // ---- queries (valid after finalize) ----
    // VDB345::get_voxel (vdb345.rs:69-106) after compute_sdf.
    int get_voxel(int32_t x, int32_t y, int32_t z, uint64_t *value, int *level) const;
    void count_nodes(uint64_t out[3]) const;
    size_t active_voxels() const;

    // Device layout (pack_layout), host copies.
    uint32_t n5() const { return (uint32_t)(n5_keys_.size() / 3); }
    uint32_t n4() const { return (uint32_t)n4_parent_.size(); }
    uint32_t n3() const { return (uint32_t)(leaf_keys_.size() / 3); }
    const std::vector<int32_t> &root_table() const { return root_; }
    const std::vector<Rec16> &n5_records() const { return n5rec_; }
    const std::vector<Rec16> &n4_records() const { return n4rec_; }
    // The leaves' layout words on the host; valid after host_words() returned true.
    const uint32_t *leaf_slots() const { return leafnib_.data(); }
    // Makes the host copy of the leaf words present (a download when a device replay
    // left them on the device only). Every query that reads leaf words calls it first.
    bool host_words(std::string &err) const;
    // The leaf words in device memory, or NULL when the replay ran on the host.
    const DeviceWords *device_words() const { return dev_words_.get(); }
    // The leaves' masks (8 words per leaf, in leaf order) that a device generator left in
    // device memory, for the device SDF replay to read instead of uploading the host copy.
    // Kept while no leaf changes: every leaf mutation, and a reordering of the leaves by
    // finalize, drops them; finalize frees them.
    void set_device_masks(std::unique_ptr<DeviceBuffer> m) {
        dev_masks_ = std::move(m);
        dev_masks_n_ = leaves_.size();
    }
    bool has_device_masks() const { return dev_masks_ != nullptr; }
    // leaf coordinates (voxel >> 3), 3 per leaf, in DFS (pool) order
    const BigVec<int32_t> &leaf_keys() const { return leaf_keys_; }

    // Reference layout, for checking against the reference's own format.
    // origins: n5 x 4; masks as u32 words (vdb345.rs:119-161); atlas slot values per node.
    void ref_masks(uint32_t *kids5, uint32_t *vals5, uint32_t *kids4, uint32_t *vals4, uint32_t *vals3) const;
    void ref_atlas(uint32_t *atlas5, uint32_t *atlas4, uint32_t *atlas3) const;

  private:
    struct Leaf {
        uint64_t mask[8];
    };
    uint32_t leaf_slot(int32_t lx, int32_t ly, int32_t lz);
    bool compute_sdf(int threads, std::string &err);
    void sdf_prepare(int threads, SdfLevelJob jobs[3]);
    uint32_t leaf_sdf(size_t leaf, uint32_t slot) const {
        return (leafnib_[leaf * 64 + (slot >> 3)] >> ((slot & 7u) * 4u)) & 15u;
    }
    bool sdf_finish(int threads, std::string &err);
    bool pack_layout(int threads, std::string &err);

    // build phase: leaf coordinate key -> pool index
    std::unordered_map<uint64_t, uint32_t> leaf_index_;
    bool index_stale_ = false;  // leaves appended by append_new_leaves are not in leaf_index_
    BigVec<int32_t> leaf_keys_;  // 3 per leaf (leaf coordinates), DFS order after finalize
    BigVec<Leaf> leaves_;
    // optional per-leaf voxel values; grown to the leaf count on the first stored value
    // (leaves past its end have none)
    std::vector<std::unique_ptr<uint32_t[]>> values_;
    // explicit internal nodes (node key -> value mask), see add_internal
    std::unordered_map<uint64_t, std::vector<uint64_t>> extra5_, extra4_;

    // topology (after finalize)
    bool finalized_ = false;
    std::vector<int32_t> n5_keys_;     // 3 per N5 (origin >> 12), sorted
    std::vector<uint32_t> n4_parent_;  // N5 index of each N4
    std::vector<uint32_t> n4_off_;     // child offset of each N4 in its N5
    std::vector<uint64_t> n5_cmask_;   // 512 per N5
    std::vector<uint64_t> n4_cmask_;   // 64 per N4
    std::vector<uint64_t> n5_vmask_;   // 512 per N5: active tiles (value mask)
    std::vector<uint64_t> n4_vmask_;   // 64 per N4
    std::vector<uint32_t> n5_first_child_;  // first N4 index per N5
    std::vector<uint32_t> n4_first_child_;  // first leaf index per N4

    // SDF values (compute_sdf), 0 for solid slots; the leaves' are kept only as the
    // packed layout words (leafnib_) once finalised
    std::vector<uint8_t> sdf5_, sdf4_;
    BigVec<uint8_t> sdf3_;
    bool leaves_packed_ = false;  // leafnib_ was written by the SDF replay itself

    // device layout
    std::vector<int32_t> root_;
    std::vector<Rec16> n5rec_, n4rec_;
    // leaf words; after a device replay filled on first use from dev_words_ (under words_mu_)
    mutable BigVec<uint32_t> leafnib_;
    std::unique_ptr<DeviceWords> dev_words_;
    mutable std::unique_ptr<std::mutex> words_mu_{new std::mutex};
    std::unique_ptr<DeviceBuffer> dev_masks_;  // see set_device_masks
    size_t dev_masks_n_ = 0;                   // the leaf count they were made for
    void drop_device_masks() {
        dev_masks_.reset();
        dev_masks_n_ = 0;
    }
};

}  // namespace wx
