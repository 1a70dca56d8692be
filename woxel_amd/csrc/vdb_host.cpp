// Host-side VDB345 builder: topology, grid SDF replay and the device layout.
// Reference semantics: src/vdb/vdb345.rs (set_voxel :26-66, get_voxel :69-106,
// masks/atlas numbering :108-264, compute_sdf :290-628) and the node index math
// of src/vdb/data_structure.rs:17-92.
#include "vdb_host.hpp"

#include "noise.hpp"

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>

namespace wx {

namespace {

constexpr int32_t kCoordLimit = 1 << 23;  // |voxel coordinate| bound of the leaf key packing
constexpr uint8_t kInf = 255;             // stands for the reference's MAX - 1 start value

inline uint64_t pack_key(int32_t a, int32_t b, int32_t c) {
    // three 21-bit biased fields; leaf coordinates lie in [-2^20, 2^20)
    const int32_t bias = 1 << 20;
    return ((uint64_t)(uint32_t)(a + bias) << 42) | ((uint64_t)(uint32_t)(b + bias) << 21) |
           (uint64_t)(uint32_t)(c + bias);
}

inline int popc64(uint64_t x) { return __builtin_popcountll(x); }

inline bool getbit(const uint64_t *m, uint32_t b) { return (m[b >> 6] >> (b & 63)) & 1ull; }
inline void setbit(uint64_t *m, uint32_t b) { m[b >> 6] |= 1ull << (b & 63); }

// Neighbour offsets of compute_sdf in the reference's push order (vdb345.rs:327-343).
struct NbSets {
    int f[13][3], b[13][3];
    NbSets() {
        int k = 0;
        for (int dy = -1; dy <= 1; ++dy)
            for (int dz = -1; dz <= 1; ++dz) {
                f[k][0] = -1, f[k][1] = dy, f[k][2] = dz;
                b[k][0] = 1, b[k][1] = dy, b[k][2] = dz;
                ++k;
            }
        for (int dz = -1; dz <= 1; ++dz) {
            f[k][0] = 0, f[k][1] = -1, f[k][2] = dz;
            b[k][0] = 0, b[k][1] = 1, b[k][2] = dz;
            ++k;
        }
        f[k][0] = 0, f[k][1] = 0, f[k][2] = -1;
        b[k][0] = 0, b[k][1] = 0, b[k][2] = 1;
    }
};
const NbSets &nb_sets() {
    static NbSets s;
    return s;
}

// Per-slot neighbour table for a node of side D = 2^LOG: for slot s and
// neighbour k, the neighbour node direction (0..26, 13 = same node) and the
// neighbour's slot offset inside that node (coordinates wrap across the border).
struct NbEntry {
    uint16_t slot;
    uint8_t dir;
    uint8_t pad;
};

template <int LOG>
const std::vector<NbEntry> &nb_table(bool forward) {
    static std::vector<NbEntry> tab[2];
    static std::once_flag once;
    std::call_once(once, [] {
        constexpr int D = 1 << LOG, S = D * D * D;
        const NbSets &ns = nb_sets();
        for (int fw = 0; fw < 2; ++fw) {
            tab[fw].resize((size_t)S * 13);
            for (int s = 0; s < S; ++s) {
                int x = s >> (2 * LOG), y = (s >> LOG) & (D - 1), z = s & (D - 1);
                for (int k = 0; k < 13; ++k) {
                    const int *dn = fw ? ns.f[k] : ns.b[k];
                    int c[3] = {x + dn[0], y + dn[1], z + dn[2]};
                    int dir[3];
                    for (int i = 0; i < 3; ++i) {
                        dir[i] = c[i] < 0 ? -1 : (c[i] >= D ? 1 : 0);
                        c[i] &= D - 1;
                    }
                    NbEntry e;
                    e.slot = (uint16_t)((c[0] << (2 * LOG)) | (c[1] << LOG) | c[2]);
                    e.dir = (uint8_t)((dir[0] + 1) * 9 + (dir[1] + 1) * 3 + (dir[2] + 1));
                    e.pad = 0;
                    tab[fw][(size_t)s * 13 + k] = e;
                }
            }
        }
    });
    return tab[forward ? 1 : 0];
}

// One node of one pass of the grid SDF. This is the per-level restatement of
// compute_sdf: a slot only ever reads slots of its own level, so the reference's
// interleaved DFS visits reduce to an order within each level.
//   in-node neighbour: solid (child node / active voxel) -> 1; empty -> min(v + 1)
//   cross-node neighbour (get_voxel): same-level empty slot -> min(v + 1); anything
//   else (no node there, a child, an active voxel, another level) -> 1
// Every candidate is >= 1, so a slot that reaches 1 is final.
template <int LOG>
inline void sdf_node(uint32_t n, const uint64_t *solid, uint8_t *val, const int32_t *nb, const NbEntry *tab,
                     bool forward) {
    constexpr uint32_t S = 1u << (3 * LOG);
    constexpr uint32_t W = S / 64;
    const int32_t *nbn = nb + (size_t)n * 27;
    const uint64_t *sn = solid + (size_t)n * W;
    uint8_t *vn = val + (size_t)n * S;
    for (uint32_t jj = 0; jj < S; ++jj) {
        const uint32_t s = forward ? jj : S - 1 - jj;
        if (getbit(sn, s)) continue;
        uint8_t v = vn[s];
        if (v == 1) continue;
        const NbEntry *e = tab + (size_t)s * 13;
        for (int k = 0; k < 13; ++k) {
            const int32_t m = nbn[e[k].dir];
            if (m < 0 || getbit(solid + (size_t)m * W, e[k].slot)) {
                v = 1;
                break;
            }
            uint8_t w = val[(size_t)m * S + e[k].slot];
            w = (w == kInf) ? kInf : (uint8_t)(w + 1);
            if (w < v) v = w;
        }
        vn[s] = v;
    }
}

// Level schedule of one pass: level(n) = 1 + max level of the neighbours the pass
// visits before n; order = nodes grouped by level, group l = order[start[l] ..
// start[l + 1]).
void sdf_schedule(uint32_t nnodes, const int32_t *nb, bool forward, std::vector<uint32_t> &order,
                  std::vector<uint32_t> &start) {
    std::vector<uint32_t> lvl(nnodes, 0);
    uint32_t maxl = 0;
    for (uint32_t ii = 0; ii < nnodes; ++ii) {
        const uint32_t n = forward ? ii : nnodes - 1 - ii;
        uint32_t l = 0;
        for (int d = 0; d < 27; ++d) {
            const int32_t m = nb[(size_t)n * 27 + d];
            if (m < 0 || (uint32_t)m == n) continue;
            if (forward ? (uint32_t)m < n : (uint32_t)m > n) l = std::max(l, lvl[(size_t)m] + 1u);
        }
        lvl[n] = l;
        maxl = std::max(maxl, l);
    }
    // bucket nodes by level (counting sort)
    start.assign(maxl + 2, 0);
    order.resize(nnodes);
    for (uint32_t n = 0; n < nnodes; ++n) start[lvl[n] + 1]++;
    for (uint32_t l = 0; l <= maxl; ++l) start[l + 1] += start[l];
    std::vector<uint32_t> fill(start.begin(), start.end() - 1);
    for (uint32_t n = 0; n < nnodes; ++n) order[fill[lvl[n]]++] = n;
}

// One pass over every node of a level. The reference visits nodes in DFS
// (= index) order, forward or reversed. A node reads only its 26 neighbour
// nodes: in the forward pass those with a smaller index must already hold
// their forward values and those with a larger index must still hold the
// start value. Level scheduling keeps exactly that: level(n) = 1 + max level of
// the earlier neighbours, every earlier neighbour sits on a lower level and every
// later neighbour on a higher one, so the nodes of one level are independent and
// run in parallel with results identical to the sequential order.
template <int LOG>
void sdf_level(uint32_t nnodes, const uint64_t *solid, uint8_t *val, const int32_t *nb, bool forward, int threads) {
    const NbEntry *tab = nb_table<LOG>(forward).data();
    if (threads <= 1 || nnodes < 256) {
        for (uint32_t ii = 0; ii < nnodes; ++ii)
            sdf_node<LOG>(forward ? ii : nnodes - 1 - ii, solid, val, nb, tab, forward);
        return;
    }
    std::vector<uint32_t> start, order;
    sdf_schedule(nnodes, nb, forward, order, start);
    const uint32_t maxl = (uint32_t)start.size() - 2;
    std::atomic<uint32_t> next{0};
    std::atomic<uint32_t> level{0};
    std::atomic<int> arrived{0};
    const int T = threads;
    auto work = [&](int) {
        for (uint32_t l = 0; l <= maxl; ++l) {
            const uint32_t b = start[l], e = start[l + 1];
            for (;;) {
                uint32_t i = next.fetch_add(16);
                if (i >= e - b) break;
                const uint32_t ie = std::min(i + 16, e - b);
                for (uint32_t k = i; k < ie; ++k) sdf_node<LOG>(order[b + k], solid, val, nb, tab, forward);
            }
            // barrier: the last thread to arrive opens the next level
            if (arrived.fetch_add(1) == T - 1) {
                next.store(0);
                arrived.store(0);
                level.store(l + 1);
            } else {
                while (level.load() <= l) std::this_thread::yield();
            }
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < T; ++t) th.emplace_back(work, t);
    work(0);
    for (auto &x : th) x.join();
}

// f(begin, end) over [0, n) split across up to `threads` threads.
template <typename F>
void parallel_for(int threads, size_t n, F f) {
    const size_t T = std::max<size_t>(1, std::min<size_t>((size_t)std::max(1, threads), n / 65536 + 1));
    if (T == 1) {
        f((size_t)0, n);
        return;
    }
    std::vector<std::thread> th;
    for (size_t t = 0; t < T; ++t) th.emplace_back(f, n * t / T, n * (t + 1) / T);
    for (auto &x : th) x.join();
}

// Sort of P = 2^k chunks in parallel, then k rounds of pairwise merges.
template <typename V>
void parallel_sort(V &v, int threads) {
    const size_t n = v.size();
    size_t P = 1;
    while (P * 2 <= (size_t)std::max(1, threads) && n / (P * 2) >= 65536) P *= 2;
    if (P == 1) {
        std::sort(v.begin(), v.end());
        return;
    }
    std::vector<size_t> cut(P + 1);
    for (size_t p = 0; p <= P; ++p) cut[p] = n * p / P;
    {
        std::vector<std::thread> th;
        for (size_t p = 0; p < P; ++p) th.emplace_back([&, p] { std::sort(v.begin() + cut[p], v.begin() + cut[p + 1]); });
        for (auto &x : th) x.join();
    }
    V tmp(n);
    for (size_t w = 1; w < P; w *= 2) {
        std::vector<std::thread> th;
        for (size_t i = 0; i < P; i += 2 * w)
            th.emplace_back([&, i, w] {
                std::merge(v.begin() + cut[i], v.begin() + cut[i + w], v.begin() + cut[i + w], v.begin() + cut[i + 2 * w],
                           tmp.begin() + cut[i]);
            });
        for (auto &x : th) x.join();
        v.swap(tmp);
    }
}

// Phase timings of a grid build on stderr when WX_TRACE_BUILD is set.
struct BuildTrace {
    bool on = std::getenv("WX_TRACE_BUILD") != nullptr;
    std::chrono::steady_clock::time_point last = std::chrono::steady_clock::now();
    void phase(const char *name) {
        if (!on) return;
        const auto t = std::chrono::steady_clock::now();
        std::fprintf(stderr, "wx build: %-22s %8.3f s\n", name, std::chrono::duration<double>(t - last).count());
        last = t;
    }
};

// Open-addressing map from packed node keys to indices (build-time lookups).
// put() may run on several threads at once for distinct keys (a slot is claimed by
// compare-and-swap); get() runs after every put() has returned.
class KeyMap {
  public:
    KeyMap(size_t n, int threads) {
        size_t cap = 16;
        while (cap < 2 * n + 16) cap <<= 1;
        keys_.resize(cap);
        vals_.resize(cap);
        parallel_for(threads, cap, [&](size_t b, size_t e) {
            for (size_t i = b; i < e; ++i) keys_[i] = ~0ull;
        });
        mask_ = cap - 1;
    }
    void put(uint64_t k, int32_t v) {
        size_t h = hash(k);
        for (;;) {
            uint64_t cur = __atomic_load_n(&keys_[h], __ATOMIC_RELAXED);
            if (cur == ~0ull) {
                if (__atomic_compare_exchange_n(&keys_[h], &cur, k, false, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
                    vals_[h] = v;
                    return;
                }
            }
            if (cur == k) {
                vals_[h] = v;
                return;
            }
            h = (h + 1) & mask_;
        }
    }
    int32_t get(uint64_t k) const {
        size_t h = hash(k);
        while (keys_[h] != ~0ull) {
            if (keys_[h] == k) return vals_[h];
            h = (h + 1) & mask_;
        }
        return -1;
    }

  private:
    size_t hash(uint64_t k) const {
        k ^= k >> 33;
        k *= 0xff51afd7ed558ccdull;
        k ^= k >> 33;
        return (size_t)k & mask_;
    }
    BigVec<uint64_t> keys_;
    BigVec<int32_t> vals_;
    size_t mask_;
};

// Neighbour-node table: nb[n*27 + d] = index of the node at key + dir(d), or -1.
template <typename Keys>
BigVec<int32_t> neighbour_nodes(const Keys &keys, int threads) {
    const size_t n = keys.size() / 3;
    KeyMap map(n, threads);
    parallel_for(threads, n, [&](size_t b, size_t e) {
        for (size_t i = b; i < e; ++i) map.put(pack_key(keys[3 * i], keys[3 * i + 1], keys[3 * i + 2]), (int32_t)i);
    });
    BigVec<int32_t> nb(n * 27);
    auto in_range = [](int32_t c) { return c >= -(1 << 20) && c < (1 << 20); };  // pack_key's domain
    auto work = [&](size_t b, size_t e) {
        for (size_t i = b; i < e; ++i)
            for (int d = 0; d < 27; ++d) {
                const int32_t x = keys[3 * i] + d / 9 - 1, y = keys[3 * i + 1] + (d / 3) % 3 - 1,
                              z = keys[3 * i + 2] + d % 3 - 1;
                nb[i * 27 + d] = in_range(x) && in_range(y) && in_range(z) ? map.get(pack_key(x, y, z)) : -1;
            }
    };
    int T = std::max(1, std::min(threads, (int)(n / 4096) + 1));
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t) th.emplace_back(work, n * t / T, n * (t + 1) / T);
    for (auto &x : th) x.join();
    return nb;
}

}  // namespace

HostVdb::HostVdb() = default;
HostVdb::~HostVdb() = default;

uint32_t HostVdb::leaf_slot(int32_t lx, int32_t ly, int32_t lz) {
    if (index_stale_) {  // leaves appended in bulk are indexed on the first lookup
        leaf_index_.clear();
        leaf_index_.reserve(leaves_.size());
        for (size_t i = 0; i < leaves_.size(); ++i)
            leaf_index_.emplace(pack_key(leaf_keys_[3 * i], leaf_keys_[3 * i + 1], leaf_keys_[3 * i + 2]), (uint32_t)i);
        index_stale_ = false;
    }
    uint64_t k = pack_key(lx, ly, lz);
    auto it = leaf_index_.find(k);
    if (it != leaf_index_.end()) return it->second;
    uint32_t id = (uint32_t)leaves_.size();
    leaf_index_.emplace(k, id);
    Leaf l;
    std::memset(l.mask, 0, sizeof l.mask);
    leaves_.push_back(l);
    leaf_keys_.push_back(lx);
    leaf_keys_.push_back(ly);
    leaf_keys_.push_back(lz);
    return id;
}

static bool coord_ok(int32_t x, int32_t y, int32_t z) {
    return x >= -kCoordLimit && x < kCoordLimit && y >= -kCoordLimit && y < kCoordLimit && z >= -kCoordLimit &&
           z < kCoordLimit;
}

bool HostVdb::set_voxel(int32_t x, int32_t y, int32_t z, uint32_t value_bits, bool store_value, std::string &err) {
    if (finalized_) { err = "set_voxel: grid already finalized"; return false; }
    drop_device_masks();
    if (!coord_ok(x, y, z)) { err = "set_voxel: coordinate outside [-2^23, 2^23)"; return false; }
    uint32_t id = leaf_slot(x >> 3, y >> 3, z >> 3);
    uint32_t o = ((uint32_t)(x & 7) << 6) | ((uint32_t)(y & 7) << 3) | (uint32_t)(z & 7);
    setbit(leaves_[id].mask, o);
    if (store_value) {
        if (values_.size() <= id) values_.resize(leaves_.size());
        if (!values_[id]) {
            values_[id].reset(new uint32_t[512]);
            std::memset(values_[id].get(), 0, 512 * 4);
        }
        values_[id][o] = value_bits;
    }
    return true;
}

bool HostVdb::add_spheres(const double *sp, size_t n, const int32_t lo[3], const int32_t hi[3], int threads,
                          std::string &err) {
    if (finalized_) { err = "add_spheres: grid already finalized"; return false; }
    drop_device_masks();
    for (int i = 0; i < 3; ++i)
        if (lo[i] < -kCoordLimit || hi[i] > kCoordLimit || lo[i] > hi[i]) {
            err = "add_spheres: clip box outside [-2^23, 2^23)";
            return false;
        }
    struct Part {
        std::vector<int32_t> keys;
        std::vector<Leaf> masks;
    };
    int T = std::max(1, threads);
    std::vector<Part> parts((size_t)T);
    std::atomic<size_t> next{0};
    auto work = [&](int t) {
        Part &P = parts[(size_t)t];
        for (;;) {
            size_t i = next.fetch_add(1);
            if (i >= n) break;
            const double c[3] = {sp[4 * i], sp[4 * i + 1], sp[4 * i + 2]};
            const double r = sp[4 * i + 3];
            const double r2 = r * r;
            int32_t a[3], b[3];
            bool empty = false;
            for (int k = 0; k < 3; ++k) {
                double fa = std::floor(c[k] - r), fb = std::ceil(c[k] + r);
                a[k] = fa < (double)lo[k] ? lo[k] : (int32_t)fa;
                b[k] = fb > (double)(hi[k] - 1) ? hi[k] - 1 : (int32_t)fb;
                if (a[k] > b[k]) empty = true;
            }
            if (empty) continue;
            for (int32_t lx = a[0] >> 3; lx <= b[0] >> 3; ++lx)
                for (int32_t ly = a[1] >> 3; ly <= b[1] >> 3; ++ly)
                    for (int32_t lz = a[2] >> 3; lz <= b[2] >> 3; ++lz) {
                        Leaf L;
                        std::memset(L.mask, 0, sizeof L.mask);
                        bool any = false;
                        const int32_t x0 = std::max(a[0], lx * 8), x1 = std::min(b[0], lx * 8 + 7);
                        const int32_t y0 = std::max(a[1], ly * 8), y1 = std::min(b[1], ly * 8 + 7);
                        const int32_t z0 = std::max(a[2], lz * 8), z1 = std::min(b[2], lz * 8 + 7);
                        for (int32_t x = x0; x <= x1; ++x) {
                            const double dx = (double)x - c[0];
                            const double dxx = dx * dx;
                            for (int32_t y = y0; y <= y1; ++y) {
                                const double dy = (double)y - c[1];
                                const double dxy = dxx + dy * dy;
                                if (dxy > r2) continue;  // adding dz*dz >= 0 cannot bring it back
                                for (int32_t z = z0; z <= z1; ++z) {
                                    const double dz = (double)z - c[2];
                                    if (dxy + dz * dz <= r2) {
                                        uint32_t o = ((uint32_t)(x & 7) << 6) | ((uint32_t)(y & 7) << 3) |
                                                     (uint32_t)(z & 7);
                                        setbit(L.mask, o);
                                        any = true;
                                    }
                                }
                            }
                        }
                        if (any) {
                            P.keys.push_back(lx);
                            P.keys.push_back(ly);
                            P.keys.push_back(lz);
                            P.masks.push_back(L);
                        }
                    }
        }
    };
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t) th.emplace_back(work, t);
    for (auto &x : th) x.join();
    for (auto &P : parts)
        if (!add_leaf_masks(P.keys.data(), P.masks.empty() ? nullptr : &P.masks.data()->mask[0], P.masks.size(), err))
            return false;
    return true;
}

bool HostVdb::noise_params(uint64_t seed, const int32_t lo[3], const int32_t hi[3], float base_freq,
                           float occupancy, wxnoise::Params &p, std::string &err) const {
    if (finalized_) { err = "add_noise: grid already finalized"; return false; }
    for (int i = 0; i < 3; ++i)
        if (lo[i] < -kCoordLimit || hi[i] > kCoordLimit || lo[i] >= hi[i]) {
            err = "add_noise: empty box or box outside [-2^23, 2^23)";
            return false;
        }
    if (!(occupancy > 0.0f && occupancy < 1.0f) || !(base_freq > 0.0f)) {
        err = "add_noise: occupancy must be in (0, 1) and base_freq > 0";
        return false;
    }
    p.seed = (uint32_t)(seed ^ (seed >> 32));
    p.f0 = base_freq;
    for (int a = 0; a < 3; ++a) p.lo[a] = lo[a], p.hi[a] = hi[a];
    // threshold = the (1 - occupancy) quantile of seeded samples
    std::vector<float> smp(65536);
    uint64_t st = seed * 0x9E3779B97F4A7C15ull + 1;
    auto next = [&] {
        st ^= st << 13;
        st ^= st >> 7;
        st ^= st << 17;
        return st;
    };
    for (float &v : smp) {
        int32_t q[3];
        for (int a = 0; a < 3; ++a) q[a] = lo[a] + (int32_t)(next() % (uint64_t)(hi[a] - lo[a]));
        v = wxnoise::noise3(p.seed, base_freq, (float)q[0], (float)q[1], (float)q[2]);
    }
    std::sort(smp.begin(), smp.end());
    p.thr = smp[(size_t)((1.0f - occupancy) * (float)(smp.size() - 1))];
    p.bound = wxnoise::leaf_bound(base_freq);
    return true;
}

bool HostVdb::add_noise(uint64_t seed, const int32_t lo[3], const int32_t hi[3], float base_freq, float occupancy,
                        int threads, float *threshold, std::string &err) {
    wxnoise::Params prm;
    if (!noise_params(seed, lo, hi, base_freq, occupancy, prm, err)) return false;
    if (threshold) *threshold = prm.thr;
    const int32_t l0[3] = {lo[0] >> 3, lo[1] >> 3, lo[2] >> 3};
    const int32_t l1[3] = {(hi[0] - 1) >> 3, (hi[1] - 1) >> 3, (hi[2] - 1) >> 3};
    const int64_t nx = l1[0] - l0[0] + 1;
    struct Part {
        std::vector<int32_t> keys;
        std::vector<Leaf> masks;
    };
    int T = std::max(1, threads);
    std::vector<Part> parts((size_t)T);
    std::atomic<int64_t> nextx{0};
    auto work = [&](int t) {
        Part &P = parts[(size_t)t];
        // lattice corners hashed once per cell (consecutive voxels share cells)
        wxnoise::Cell cache[wxnoise::kOctaves];
        wxnoise::clear_cache(cache);
        for (;;) {
            const int64_t ix = nextx.fetch_add(1);
            if (ix >= nx) break;
            const int32_t lx = l0[0] + (int32_t)ix;
            for (int32_t ly = l0[1]; ly <= l1[1]; ++ly)
                for (int32_t lz = l0[2]; lz <= l1[2]; ++lz) {
                    const int32_t b[3] = {lx * 8, ly * 8, lz * 8};
                    bool inside_box;
                    const int cls = wxnoise::leaf_class(prm, b, inside_box);
                    if (cls == 0) continue;  // empty leaf
                    Leaf L;
                    std::memset(L.mask, 0, sizeof L.mask);
                    bool any = false;
                    if (cls == 1) {
                        std::memset(L.mask, 0xFF, sizeof L.mask);
                        any = true;
                    } else {
                        for (int x = 0; x < 8; ++x)
                            for (int y = 0; y < 8; ++y)
                                for (int z = 0; z < 8; ++z)
                                    if (wxnoise::voxel_active(prm, b[0] + x, b[1] + y, b[2] + z, cache)) {
                                        setbit(L.mask, (uint32_t)(x << 6 | y << 3 | z));
                                        any = true;
                                    }
                    }
                    if (any) {
                        P.keys.push_back(lx);
                        P.keys.push_back(ly);
                        P.keys.push_back(lz);
                        P.masks.push_back(L);
                    }
                }
        }
    };
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t) th.emplace_back(work, t);
    for (auto &x : th) x.join();
    for (auto &P : parts)
        if (!add_leaf_masks(P.keys.data(), P.masks.empty() ? nullptr : &P.masks.data()->mask[0], P.masks.size(), err))
            return false;
    return true;
}

bool HostVdb::append_new_leaves(const int32_t *keys, const uint64_t *masks, size_t n, std::string &err) {
    if (finalized_) { err = "add_leaf_masks: grid already finalized"; return false; }
    int32_t *k = nullptr;
    uint64_t *m = nullptr;
    if (!grow_new_leaves(n, &k, &m, err)) return false;
    std::memcpy(m, masks, n * sizeof(Leaf));
    std::memcpy(k, keys, n * 12);
    return true;
}

bool HostVdb::grow_new_leaves(size_t n, int32_t **keys, uint64_t **masks, std::string &err) {
    if (finalized_) { err = "add_leaf_masks: grid already finalized"; return false; }
    drop_device_masks();
    const size_t base = leaves_.size();
    leaves_.resize(base + n);  // BigVec: no value initialisation
    leaf_keys_.resize(3 * (base + n));
    *masks = n ? leaves_[base].mask : nullptr;
    *keys = n ? &leaf_keys_[3 * base] : nullptr;
    index_stale_ = index_stale_ || n > 0;
    return true;
}

void HostVdb::drop_leaves_from(size_t base) {
    if (finalized_ || base >= leaves_.size()) return;
    drop_device_masks();
    leaves_.resize(base);
    leaf_keys_.resize(3 * base);
    if (values_.size() > base) values_.resize(base);
    index_stale_ = true;
}

bool HostVdb::add_leaf_masks(const int32_t *keys, const uint64_t *masks, size_t n, std::string &err,
                             bool keep_empty) {
    if (finalized_) { err = "add_leaf_masks: grid already finalized"; return false; }
    drop_device_masks();
    for (size_t i = 0; i < n; ++i) {
        const int32_t *k = keys + 3 * i;
        if (!coord_ok(k[0] * 8, k[1] * 8, k[2] * 8) || k[0] >= (1 << 20) || k[1] >= (1 << 20) || k[2] >= (1 << 20)) {
            err = "add_leaf_masks: leaf key outside [-2^20, 2^20)";
            return false;
        }
        const uint64_t *m = masks + 8 * i;
        uint64_t any = 0;
        for (int w = 0; w < 8; ++w) any |= m[w];
        if (!any && !keep_empty) continue;  // set_voxel only makes a leaf for a set voxel
        uint32_t id = leaf_slot(k[0], k[1], k[2]);
        for (int w = 0; w < 8; ++w) leaves_[id].mask[w] |= m[w];
    }
    return true;
}

bool HostVdb::add_internal(int level, const int32_t key[3], const uint64_t *value_mask, std::string &err) {
    if (finalized_) { err = "add_internal: grid already finalized"; return false; }
    if (level != 5 && level != 4) { err = "add_internal: level must be 5 or 4"; return false; }
    const int shift = level == 5 ? 12 : 7;
    for (int a = 0; a < 3; ++a)
        if (key[a] < -(kCoordLimit >> shift) || key[a] >= (kCoordLimit >> shift)) {
            err = "add_internal: node outside [-2^23, 2^23)";
            return false;
        }
    const size_t words = level == 5 ? 512 : 64;
    auto &m = (level == 5 ? extra5_ : extra4_)[pack_key(key[0], key[1], key[2])];
    if (m.empty()) m.assign(words, 0);
    if (value_mask)
        for (size_t w = 0; w < words; ++w) m[w] |= value_mask[w];
    return true;
}

void HostVdb::sdf_prepare(int threads, SdfLevelJob jobs[3]) {
    const size_t n5c = n5_keys_.size() / 3, n4c = n4_parent_.size(), nl = leaves_.size();
    std::vector<int32_t> k4(3 * n4c);
    for (size_t j = 0; j < n4c; ++j) {
        const int32_t *p = &n5_keys_[3 * (size_t)n4_parent_[j]];
        uint32_t o = n4_off_[j];
        k4[3 * j] = p[0] * 32 + (int32_t)(o >> 10);
        k4[3 * j + 1] = p[1] * 32 + (int32_t)((o >> 5) & 31);
        k4[3 * j + 2] = p[2] * 32 + (int32_t)(o & 31);
    }
    BuildTrace trace;
    sdf5_.assign(n5c * 32768, kInf);
    sdf4_.assign(n4c * 4096, kInf);
    leaves_packed_ = true;  // the replay produces the leaf words (kept on the device)
    const int logs[3] = {5, 4, 3};
    const size_t counts[3] = {n5c, n4c, nl};
    const uint64_t *solids[3] = {n5_cmask_.data(), n4_cmask_.data(), nl ? leaves_[0].mask : nullptr};
    uint8_t *vals[3] = {sdf5_.data(), sdf4_.data(), nullptr};
    for (int l = 0; l < 3; ++l) {
        SdfLevelJob &J = jobs[l];
        J.log = logs[l];
        J.nodes = (uint32_t)counts[l];
        J.solid = solids[l];
        J.val = vals[l];
        if (l == 2) {  // the leaves' tables are built by the device replay
            J.keys = leaf_keys_.data();
            if (dev_masks_ && dev_masks_n_ == nl) {
                J.solid_dev = dev_masks_->data();
                J.solid_dev_device = dev_masks_->device();
            }
            break;
        }
        J.nb = neighbour_nodes(l == 0 ? n5_keys_ : k4, threads);
        trace.phase("  node neighbour table");
        // the two schedules are independent walks over the same table
        std::thread bw([&J] { sdf_schedule(J.nodes, J.nb.data(), false, J.order[0], J.start[0]); });
        sdf_schedule(J.nodes, J.nb.data(), true, J.order[1], J.start[1]);
        bw.join();
        trace.phase("  node schedules");
    }
    jobs[2].words = true;
}

bool HostVdb::finalize(int threads, std::string &err) { return finalize_with(threads, nullptr, nullptr, err); }

bool HostVdb::host_words(std::string &err) const {
    const size_t want = leaves_.size() * 64;
    std::lock_guard<std::mutex> lk(*words_mu_);
    if (leafnib_.size() == want || !dev_words_) return true;
    leafnib_.resize(want);
    if (!dev_words_->download(leafnib_.data(), want, err)) {
        BigVec<uint32_t>().swap(leafnib_);
        return false;
    }
    return true;
}

bool HostVdb::finalize_with(int threads, SdfRunner run, void *ctx, std::string &err) {
    if (finalized_) return true;
    const size_t nl = leaves_.size();
    BuildTrace trace;
    // DFS order of leaves: N5 key (lexicographic, signed), then N5 child offset,
    // then N4 child offset -- the numbering of vdb345.rs:108-264 -- as one 63-bit
    // key: the N5 key fields biased to 12 bits each, off4 (15 bits), off3 (12 bits)
    auto dfs_key = [&](size_t i) {
        const int32_t *k = &leaf_keys_[3 * i];
        const uint64_t off4 = ((uint64_t)((k[0] >> 4) & 31) << 10) | ((uint64_t)((k[1] >> 4) & 31) << 5) |
                              (uint64_t)((k[2] >> 4) & 31);
        const uint64_t off3 = ((uint64_t)(k[0] & 15) << 8) | ((uint64_t)(k[1] & 15) << 4) | (uint64_t)(k[2] & 15);
        return ((uint64_t)((k[0] >> 9) + 2048) << 51) | ((uint64_t)((k[1] >> 9) + 2048) << 39) |
               ((uint64_t)((k[2] >> 9) + 2048) << 27) | (off4 << 12) | off3;
    };
    // leaves that arrive in DFS order (the device generator sorts its output) keep
    // their pool positions: no key array, no sort and no permutation (leaf keys are unique)
    std::atomic<bool> sorted{true};
    parallel_for(threads, nl, [&](size_t b, size_t e) {
        for (size_t i = b; i < e && i + 1 < nl; ++i)
            if (dfs_key(i + 1) < dfs_key(i)) {
                sorted.store(false, std::memory_order_relaxed);
                return;
            }
    });
    if (sorted.load()) {
        trace.phase("DFS sort");
        leaf_index_.clear();
        index_stale_ = false;
    } else {  // permute the leaf pool into DFS order
        BigVec<std::pair<uint64_t, uint32_t>> sk(nl);  // filled (first touched) by the threads below
        parallel_for(threads, nl, [&](size_t b, size_t e) {
            for (size_t i = b; i < e; ++i) sk[i] = {dfs_key(i), (uint32_t)i};
        });
        parallel_sort(sk, threads);
        trace.phase("DFS sort");
        BigVec<Leaf> L(nl);
        BigVec<int32_t> K(3 * nl);
        std::vector<std::unique_ptr<uint32_t[]>> V(values_.empty() ? 0 : nl);
        parallel_for(threads, nl, [&](size_t b, size_t e) {
            for (size_t i = b; i < e; ++i) {
                L[i] = leaves_[sk[i].second];
                std::memcpy(&K[3 * i], &leaf_keys_[3 * (size_t)sk[i].second], 12);
                if (!V.empty() && sk[i].second < values_.size()) V[i] = std::move(values_[sk[i].second]);
            }
        });
        leaves_.swap(L);
        drop_device_masks();  // they follow the old order
        leaf_keys_.swap(K);
        values_.swap(V);
        leaf_index_.clear();
        index_stale_ = false;
    }
    auto k5_of = [](uint64_t key, int32_t out[3]) {
        for (int a = 0; a < 3; ++a) out[a] = (int32_t)((key >> (51 - 12 * a)) & 0xFFFu) - 2048;
    };
    trace.phase("leaf permutation");
    // N5 / N4 node lists: the ancestors of the leaves and the explicit nodes, in
    // DFS order (N5 key, then child offset), and the child masks
    struct N4Key {
        int32_t k5[3];
        uint32_t off4;
        bool operator<(const N4Key &b) const {
            for (int i = 0; i < 3; ++i)
                if (k5[i] != b.k5[i]) return k5[i] < b.k5[i];
            return off4 < b.off4;
        }
        bool operator==(const N4Key &b) const {
            return k5[0] == b.k5[0] && k5[1] == b.k5[1] && k5[2] == b.k5[2] && off4 == b.off4;
        }
    };
    auto unpack = [](uint64_t k, int32_t out[3]) {
        for (int a = 0; a < 3; ++a) {
            const uint32_t f = (uint32_t)(k >> (21 * (2 - a))) & 0x1FFFFFu;
            out[a] = (int32_t)f - (1 << 20);  // pack_key's bias
        }
    };
    auto n4key_of = [](const int32_t k4[3]) {
        N4Key q;
        for (int a = 0; a < 3; ++a) q.k5[a] = k4[a] >> 5;
        q.off4 = ((uint32_t)(k4[0] & 31) << 10) | ((uint32_t)(k4[1] & 31) << 5) | (uint32_t)(k4[2] & 31);
        return q;
    };
    // the leaves' N4s: the first leaf of each run of equal sk >> 12, collected per chunk of
    // leaves on its own thread and joined in chunk order
    std::vector<N4Key> c4;
    {
        const size_t T = std::max<size_t>(1, std::min<size_t>((size_t)std::max(1, threads), nl / 65536 + 1));
        std::vector<std::vector<N4Key>> part(T);
        std::vector<std::thread> th;
        for (size_t t = 0; t < T; ++t)
            th.emplace_back([&, t] {
                uint64_t prev = nl * t / T ? dfs_key(nl * t / T - 1) >> 12 : ~0ull;
                for (size_t i = nl * t / T, e = nl * (t + 1) / T; i < e; ++i) {
                    const uint64_t key = dfs_key(i);
                    if ((key >> 12) == prev) continue;  // same N4
                    prev = key >> 12;
                    N4Key q;
                    k5_of(key, q.k5);
                    q.off4 = (uint32_t)(key >> 12) & 0x7FFFu;
                    part[t].push_back(q);
                }
            });
        for (auto &x : th) x.join();
        size_t total = extra4_.size();
        for (const auto &v : part) total += v.size();
        c4.reserve(total);
        for (const auto &v : part) c4.insert(c4.end(), v.begin(), v.end());
    }
    for (const auto &e : extra4_) {
        int32_t k4[3];
        unpack(e.first, k4);
        c4.push_back(n4key_of(k4));
    }
    std::sort(c4.begin(), c4.end());
    c4.erase(std::unique(c4.begin(), c4.end()), c4.end());
    std::vector<std::array<int32_t, 3>> c5;
    c5.reserve(c4.size() + extra5_.size());
    for (const N4Key &q : c4) c5.push_back({q.k5[0], q.k5[1], q.k5[2]});
    for (const auto &e : extra5_) {
        std::array<int32_t, 3> k;
        unpack(e.first, k.data());
        c5.push_back(k);
    }
    std::sort(c5.begin(), c5.end());
    c5.erase(std::unique(c5.begin(), c5.end()), c5.end());
    const size_t n5c = c5.size(), n4c = c4.size();
    n5_keys_.resize(3 * n5c);
    for (size_t i = 0; i < n5c; ++i) std::memcpy(&n5_keys_[3 * i], c5[i].data(), 12);
    n4_parent_.resize(n4c);
    n4_off_.resize(n4c);
    n5_first_child_.assign(n5c, 0);
    for (size_t j = 0, i5 = 0; i5 < n5c; ++i5) {  // both lists are sorted by N5 key
        n5_first_child_[i5] = (uint32_t)j;
        while (j < n4c && std::memcmp(c4[j].k5, c5[i5].data(), 12) == 0) {
            n4_parent_[j] = (uint32_t)i5;
            n4_off_[j] = c4[j].off4;
            ++j;
        }
    }
    // leaves are sorted by (N5 key, off4, off3) = sk: N4 j's leaves start at the first leaf
    // whose sk >> 12 is not below j's key (an N4 without leaves gets an empty range)
    n4_first_child_.assign(n4c, 0);
    n5_cmask_.assign(n5c * 512, 0);
    n4_cmask_.assign(n4c * 64, 0);
    n5_vmask_.assign(n5c * 512, 0);
    n4_vmask_.assign(n4c * 64, 0);
    auto key4 = [](const N4Key &q) {
        return ((uint64_t)(q.k5[0] + 2048) << 39) | ((uint64_t)(q.k5[1] + 2048) << 27) |
               ((uint64_t)(q.k5[2] + 2048) << 15) | (uint64_t)q.off4;
    };
    parallel_for(threads, n4c, [&](size_t b, size_t e) {
        for (size_t j = b; j < e; ++j) {
            const uint64_t k = key4(c4[j]);
            size_t lo = 0, hi = nl;  // the first leaf whose N4 key is not below k
            while (lo < hi) {
                const size_t mid = lo + (hi - lo) / 2;
                if ((dfs_key(mid) >> 12) < k) lo = mid + 1;
                else hi = mid;
            }
            n4_first_child_[j] = (uint32_t)lo;
        }
    });
    // each N4's child mask from its own leaves (one thread per range of N4s: an N4 holds up
    // to 4096 leaves, so the split is by N4 count whatever the total)
    {
        const size_t T = std::max<size_t>(1, std::min<size_t>((size_t)std::max(1, threads), nl / 65536 + 1));
        std::vector<std::thread> th;
        for (size_t t = 0; t < T; ++t)
            th.emplace_back([&, t] {
                for (size_t j = n4c * t / T, je = n4c * (t + 1) / T; j < je; ++j) {
                    const size_t end = j + 1 < n4c ? n4_first_child_[j + 1] : nl;
                    for (size_t i = n4_first_child_[j]; i < end; ++i)
                        setbit(&n4_cmask_[j * 64], (uint32_t)dfs_key(i) & 0xFFFu);
                }
            });
        for (auto &x : th) x.join();
    }
    for (size_t j = 0; j < n4c; ++j) setbit(&n5_cmask_[(size_t)n4_parent_[j] * 512], n4_off_[j]);
    for (const auto &e : extra5_) {
        std::array<int32_t, 3> k;
        unpack(e.first, k.data());
        const size_t i5 = (size_t)(std::lower_bound(c5.begin(), c5.end(), k) - c5.begin());
        std::memcpy(&n5_vmask_[i5 * 512], e.second.data(), 512 * 8);
    }
    for (const auto &e : extra4_) {
        int32_t k4[3];
        unpack(e.first, k4);
        const size_t j = (size_t)(std::lower_bound(c4.begin(), c4.end(), n4key_of(k4)) - c4.begin());
        std::memcpy(&n4_vmask_[j * 64], e.second.data(), 64 * 8);
    }
    extra5_.clear();
    extra4_.clear();
    trace.phase("topology");
    if (run) {
        SdfLevelJob jobs[3];
        sdf_prepare(threads, jobs);
        trace.phase("SDF schedules");
        if (!run(jobs, ctx, err) || !sdf_finish(threads, err)) return false;
        if (nl && !jobs[2].kept) {
            err = "compute_sdf: the device replay returned no leaf words";
            return false;
        }
        dev_words_ = std::move(jobs[2].kept);
        BigVec<uint32_t>().swap(leafnib_);
        trace.phase("SDF replay (device)");
    } else {
        if (!compute_sdf(threads, err)) return false;
        trace.phase("SDF replay (host)");
    }
    drop_device_masks();  // read by the replay only
    if (!pack_layout(threads, err)) return false;
    trace.phase("layout");
    finalized_ = true;
    return true;
}

bool HostVdb::compute_sdf(int threads, std::string &err) {
    const size_t n5c = n5_keys_.size() / 3, n4c = n4_parent_.size(), nl = leaves_.size();
    // node keys in each level's own units
    std::vector<int32_t> k4(3 * n4c);
    for (size_t j = 0; j < n4c; ++j) {
        const int32_t *p = &n5_keys_[3 * (size_t)n4_parent_[j]];
        uint32_t o = n4_off_[j];
        k4[3 * j] = p[0] * 32 + (int32_t)(o >> 10);
        k4[3 * j + 1] = p[1] * 32 + (int32_t)((o >> 5) & 31);
        k4[3 * j + 2] = p[2] * 32 + (int32_t)(o & 31);
    }
    BigVec<int32_t> nb5 = neighbour_nodes(n5_keys_, threads);
    BigVec<int32_t> nb4 = neighbour_nodes(k4, threads);
    BigVec<int32_t> nb3 = neighbour_nodes(leaf_keys_, threads);
    sdf5_.assign(n5c * 32768, kInf);
    sdf4_.assign(n4c * 4096, kInf);
    sdf3_.resize(nl * 512);
    parallel_for(threads, sdf3_.size(), [&](size_t b, size_t e) { std::memset(sdf3_.data() + b, kInf, e - b); });
    // leaf solid mask = active voxels (the leaf pool is 8 words per leaf)
    const uint64_t *solid3 = nl ? leaves_[0].mask : nullptr;
    // the three levels are independent: the two upper levels run beside the leaves
    auto lvl5 = [&] {
        sdf_level<5>((uint32_t)n5c, n5_cmask_.data(), sdf5_.data(), nb5.data(), true, 1);
        sdf_level<5>((uint32_t)n5c, n5_cmask_.data(), sdf5_.data(), nb5.data(), false, 1);
    };
    auto lvl4 = [&] {
        sdf_level<4>((uint32_t)n4c, n4_cmask_.data(), sdf4_.data(), nb4.data(), true, 1);
        sdf_level<4>((uint32_t)n4c, n4_cmask_.data(), sdf4_.data(), nb4.data(), false, 1);
    };
    const int t3 = std::max(1, threads - 1);
    auto lvl3 = [&] {
        sdf_level<3>((uint32_t)nl, solid3, sdf3_.data(), nb3.data(), true, t3);
        sdf_level<3>((uint32_t)nl, solid3, sdf3_.data(), nb3.data(), false, t3);
    };
    if (threads > 1) {
        std::thread t54([&] {
            lvl5();
            lvl4();
        });
        lvl3();
        t54.join();
    } else {
        lvl5();
        lvl4();
        lvl3();
    }
    return sdf_finish(threads, err);
}

// Zero the solid slots and check the value ranges of the device layout.
bool HostVdb::sdf_finish(int threads, std::string &err) {
    const size_t n5c = n5_keys_.size() / 3, n4c = n4_parent_.size(), nl = leaves_.size();
    for (size_t n = 0; n < n5c; ++n)
        for (uint32_t s = 0; s < 32768; ++s) {
            uint8_t &v = sdf5_[n * 32768 + s];
            if (getbit(&n5_cmask_[n * 512], s)) v = 0;
            else if (v == kInf) { err = "compute_sdf: node-5 tile without a finite distance (unsupported topology)"; return false; }
        }
    for (size_t n = 0; n < n4c; ++n)
        for (uint32_t s = 0; s < 4096; ++s) {
            uint8_t &v = sdf4_[n * 4096 + s];
            if (getbit(&n4_cmask_[n * 64], s)) v = 0;
            else if (v > 15) { err = "compute_sdf: node-4 tile distance above 15 (unsupported topology)"; return false; }
        }
    if (leaves_packed_) return true;  // the replay range-checked the leaves itself
    std::atomic<bool> over{false};
    parallel_for(threads, nl, [&](size_t b, size_t e) {
        for (size_t n = b; n < e; ++n)
            for (uint32_t s = 0; s < 512; ++s) {
                uint8_t &v = sdf3_[n * 512 + s];
                if (getbit(leaves_[n].mask, s)) v = 0;
                else if (v > 15) over.store(true, std::memory_order_relaxed);
            }
    });
    if (over.load()) { err = "compute_sdf: voxel distance above 15 (unsupported topology)"; return false; }
    return true;
}

bool HostVdb::pack_layout(int threads, std::string &err) {
    (void)err;
    const size_t n5c = n5_keys_.size() / 3, n4c = n4_parent_.size(), nl = leaves_.size();
    root_.assign(n5c * 4, 0);
    for (size_t n = 0; n < n5c; ++n) {
        std::memcpy(&root_[4 * n], &n5_keys_[3 * n], 12);
        root_[4 * n + 3] = (int32_t)n;
    }
    // node-5 records: 8 slots each = {kids8 | walk8 << 8 | vals8 << 16, first child index, sdf bytes 0-3,
    // 4-7}; walk = a child that is not an active tile (the lookup descends there), and an active
    // tile's byte is 0, so the lookup reads a hit from the byte alone (vals is kept for decoding)
    n5rec_.assign(n5c * 4096, Rec16{{0, 0, 0, 0}});
    for (size_t n = 0; n < n5c; ++n) {
        uint32_t base = n5_first_child_[n];
        const uint64_t *cm = &n5_cmask_[n * 512];
        const uint64_t *vm = &n5_vmask_[n * 512];
        for (uint32_t r = 0; r < 4096; ++r) {
            uint32_t kids = (uint32_t)(cm[r >> 3] >> ((r & 7) * 8)) & 0xFFu;
            uint32_t vals = (uint32_t)(vm[r >> 3] >> ((r & 7) * 8)) & 0xFFu;
            Rec16 &R = n5rec_[n * 4096 + r];
            R.w[0] = kids | ((kids & ~vals) << 8) | (vals << 16);
            R.w[1] = base;
            const uint8_t *sd = &sdf5_[n * 32768 + (size_t)r * 8];
            uint32_t b8[8];
            for (int k = 0; k < 8; ++k) b8[k] = ((vals >> k) & 1u) ? 0u : (uint32_t)sd[k];
            R.w[2] = b8[0] | (b8[1] << 8) | (b8[2] << 16) | (b8[3] << 24);
            R.w[3] = b8[4] | (b8[5] << 8) | (b8[6] << 16) | (b8[7] << 24);
            base += (uint32_t)popc64(kids);
        }
    }
    // node-4 records: 16 slots each = {kids16 | walk16 << 16, first child index, nibbles 0-7, 8-15};
    // an active tile's nibble is 0 (inactive tiles are >= 1), which is how vals16 is recovered
    n4rec_.assign(n4c * 256, Rec16{{0, 0, 0, 0}});
    for (size_t n = 0; n < n4c; ++n) {
        uint32_t base = n4_first_child_[n];
        const uint64_t *cm = &n4_cmask_[n * 64];
        const uint64_t *vm = &n4_vmask_[n * 64];
        for (uint32_t r = 0; r < 256; ++r) {
            uint32_t kids = (uint32_t)(cm[r >> 2] >> ((r & 3) * 16)) & 0xFFFFu;
            uint32_t vals = (uint32_t)(vm[r >> 2] >> ((r & 3) * 16)) & 0xFFFFu;
            Rec16 &R = n4rec_[n * 256 + r];
            R.w[0] = kids | ((kids & ~vals) << 16);
            R.w[1] = base;
            const uint8_t *sd = &sdf4_[n * 4096 + (size_t)r * 16];
            uint32_t lo = 0, hi = 0;
            for (int k = 0; k < 8; ++k) {
                lo |= (((vals >> k) & 1u) ? 0u : (uint32_t)sd[k]) << (4 * k);
                hi |= (((vals >> (8 + k)) & 1u) ? 0u : (uint32_t)sd[8 + k]) << (4 * k);
            }
            R.w[2] = lo;
            R.w[3] = hi;
            base += (uint32_t)popc64(kids);
        }
    }
    // leaves: 64 words of 8 nibbles; nibble 0 = active voxel, else the SDF distance
    if (!leaves_packed_) {
        leafnib_.resize(nl * 64);
        parallel_for(threads, nl, [&](size_t b, size_t e) {
            for (size_t n = b; n < e; ++n)
                for (uint32_t w = 0; w < 64; ++w) {
                    uint32_t word = 0;
                    for (int k = 0; k < 8; ++k) word |= (uint32_t)sdf3_[n * 512 + w * 8 + k] << (4 * k);
                    leafnib_[n * 64 + w] = word;
                }
        });
        leaves_packed_ = true;
    }
    BigVec<uint8_t>().swap(sdf3_);  // the words hold the same values
    return true;
}

int HostVdb::get_voxel(int32_t x, int32_t y, int32_t z, uint64_t *value, int *level) const {
    *level = 0;
    *value = 0;
    const int32_t k5[3] = {x >> 12, y >> 12, z >> 12};
    const size_t n5c = n5_keys_.size() / 3;
    size_t i5 = n5c;
    for (size_t n = 0; n < n5c; ++n)
        if (std::memcmp(&n5_keys_[3 * n], k5, 12) == 0) { i5 = n; break; }
    if (i5 == n5c) return EP_BKGR;
    const uint32_t o5 = ((uint32_t)((x >> 7) & 31) << 10) | ((uint32_t)((y >> 7) & 31) << 5) | (uint32_t)((z >> 7) & 31);
    if (!getbit(&n5_cmask_[i5 * 512], o5)) {
        *value = sdf5_[i5 * 32768 + o5];
        *level = 5;
        return EP_INNR;
    }
    const Rec16 &R5 = n5rec_[i5 * 4096 + (o5 >> 3)];
    uint32_t i4 = R5.w[1] + (uint32_t)__builtin_popcount((R5.w[0] & 0xFFu) & ((1u << (o5 & 7)) - 1u));
    const uint32_t o4 = ((uint32_t)((x >> 3) & 15) << 8) | ((uint32_t)((y >> 3) & 15) << 4) | (uint32_t)((z >> 3) & 15);
    if (!getbit(&n4_cmask_[(size_t)i4 * 64], o4)) {
        *value = sdf4_[(size_t)i4 * 4096 + o4];
        *level = 4;
        return EP_INNR;
    }
    const Rec16 &R4 = n4rec_[(size_t)i4 * 256 + (o4 >> 4)];
    uint32_t i3 = R4.w[1] + (uint32_t)__builtin_popcount((R4.w[0] & 0xFFFFu) & ((1u << (o4 & 15)) - 1u));
    const uint32_t o3 = ((uint32_t)(x & 7) << 6) | ((uint32_t)(y & 7) << 3) | (uint32_t)(z & 7);
    if (getbit(leaves_[i3].mask, o3)) {
        *value = i3 < values_.size() && values_[i3] ? values_[i3][o3] : 0;
        return EP_LEAF;
    }
    *value = leaf_sdf(i3, o3);
    return EP_OFFS;
}

void HostVdb::count_nodes(uint64_t out[3]) const {
    out[0] = n5_keys_.size() / 3;
    out[1] = n4_parent_.size();
    out[2] = leaves_.size();
}

size_t HostVdb::active_voxels() const {
    size_t s = 0;
    for (const Leaf &l : leaves_)
        for (int w = 0; w < 8; ++w) s += (size_t)popc64(l.mask[w]);
    return s;
}

void HostVdb::ref_masks(uint32_t *kids5, uint32_t *vals5, uint32_t *kids4, uint32_t *vals4, uint32_t *vals3) const {
    auto split = [](const uint64_t *a, size_t n, uint32_t *o) {
        for (size_t i = 0; i < n; ++i) {
            o[2 * i] = (uint32_t)a[i];
            o[2 * i + 1] = (uint32_t)(a[i] >> 32);
        }
    };
    split(n5_cmask_.data(), n5_cmask_.size(), kids5);
    split(n5_vmask_.data(), n5_vmask_.size(), vals5);
    split(n4_cmask_.data(), n4_cmask_.size(), kids4);
    split(n4_vmask_.data(), n4_vmask_.size(), vals4);
    for (size_t i = 0; i < leaves_.size(); ++i) split(leaves_[i].mask, 8, vals3 + 16 * i);
}

void HostVdb::ref_atlas(uint32_t *atlas5, uint32_t *atlas4, uint32_t *atlas3) const {
    const size_t n5c = n5_keys_.size() / 3, n4c = n4_parent_.size(), nl = leaves_.size();
    for (size_t n = 0; n < n5c; ++n) {
        uint32_t c = n5_first_child_[n];
        for (uint32_t s = 0; s < 32768; ++s)
            atlas5[n * 32768 + s] = getbit(&n5_cmask_[n * 512], s) ? c++ : sdf5_[n * 32768 + s];
    }
    for (size_t n = 0; n < n4c; ++n) {
        uint32_t c = n4_first_child_[n];
        for (uint32_t s = 0; s < 4096; ++s)
            atlas4[n * 4096 + s] = getbit(&n4_cmask_[n * 64], s) ? c++ : sdf4_[n * 4096 + s];
    }
    for (size_t n = 0; n < nl; ++n)
        for (uint32_t s = 0; s < 512; ++s)
            atlas3[n * 512 + s] =
                getbit(leaves_[n].mask, s) ? (n < values_.size() && values_[n] ? values_[n][s] : 0u) : leaf_sdf(n, s);
}

}  // namespace wx
