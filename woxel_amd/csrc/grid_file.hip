// Grid-layout cache: one finalized device layout per file, written and read
// without a host copy of the pools.
//
// The reference rebuilds its tree (and compute_sdf) from the scene every run; its
// only serialised form is the .vdb writer (src/vdb/write.rs:10-41), which stores
// topology and values, not the SDF. A cached layout here holds the root table, the
// node records with their tile distances and the leaf words, so a grid that took
// seconds to build comes back at disk (or page-cache) speed, bit for bit.
//
// File (little-endian):
//   header, kHeaderBytes:
//     0    "WXLAYOUT"                8    u32 format (1)       12   u32 header bytes
//     16   u32 n5, n4, n3            28   f32 bound            32   u32 key length
//     36   u32 chunk bytes           40   u64 offset[4]        72   u64 bytes[4]
//     104  u64 hash[4]               256  key bytes            4088 u64 XXH64 of bytes [0, 4088)
//   sections root, node-5 records, node-4 records, leaf words, each at a 4 KiB-aligned offset.
// A section's hash is XXH64 (seed 0) over the little-endian XXH64 digests (seed 0) of
// its consecutive chunk-byte pieces, so pieces hash in parallel while they stream.
// XXH64 is Yann Collet's published 64-bit xxHash (BSD-2), restated from its
// specification.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "lane.hpp"
#include "wx_internal.hpp"

using wx::set_error;

namespace {

constexpr char kMagic[8] = {'W', 'X', 'L', 'A', 'Y', 'O', 'U', 'T'};
constexpr uint32_t kFormat = 1u;
constexpr uint32_t kHeaderBytes = 4096u;
constexpr uint32_t kKeyOff = 256u, kKeyMax = 2048u, kHashOff = kHeaderBytes - 8u;
constexpr uint64_t kChunk = 16ull << 20;  // hashing piece
constexpr uint64_t kStage = 64ull << 20;  // staging buffer: 4 pieces

// ---- XXH64 ----
constexpr uint64_t P1 = 0x9E3779B185EBCA87ull, P2 = 0xC2B2AE3D27D4EB4Full, P3 = 0x165667B19E3779F9ull,
                   P4 = 0x85EBCA77C2B2AE63ull, P5 = 0x27D4EB2F165667C5ull;

inline uint64_t rotl(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }
inline uint64_t rd64(const unsigned char *p) {
    uint64_t v;
    std::memcpy(&v, p, 8);
    return v;
}
inline uint32_t rd32(const unsigned char *p) {
    uint32_t v;
    std::memcpy(&v, p, 4);
    return v;
}
inline uint64_t xround(uint64_t acc, uint64_t in) { return rotl(acc + in * P2, 31) * P1; }
inline uint64_t xmerge(uint64_t h, uint64_t v) { return (h ^ xround(0, v)) * P1 + P4; }

uint64_t xxh64(const void *data, size_t len, uint64_t seed) {
    const unsigned char *p = static_cast<const unsigned char *>(data), *end = p + len;
    uint64_t h;
    if (len >= 32) {
        uint64_t v1 = seed + P1 + P2, v2 = seed + P2, v3 = seed, v4 = seed - P1;
        const unsigned char *lim = end - 32;
        do {
            v1 = xround(v1, rd64(p));
            v2 = xround(v2, rd64(p + 8));
            v3 = xround(v3, rd64(p + 16));
            v4 = xround(v4, rd64(p + 24));
            p += 32;
        } while (p <= lim);
        h = rotl(v1, 1) + rotl(v2, 7) + rotl(v3, 12) + rotl(v4, 18);
        h = xmerge(h, v1);
        h = xmerge(h, v2);
        h = xmerge(h, v3);
        h = xmerge(h, v4);
    } else {
        h = seed + P5;
    }
    h += (uint64_t)len;
    for (; p + 8 <= end; p += 8) h = rotl(h ^ xround(0, rd64(p)), 27) * P1 + P4;
    if (p + 4 <= end) {
        h = rotl(h ^ (uint64_t)rd32(p) * P1, 23) * P2 + P3;
        p += 4;
    }
    for (; p < end; ++p) h = rotl(h ^ (uint64_t)*p * P5, 11) * P1;
    h ^= h >> 33;
    h *= P2;
    h ^= h >> 29;
    h *= P3;
    h ^= h >> 32;
    return h;
}

// Digests of the kChunk pieces of buf[0, n) (n a multiple of kChunk except at the
// section's end), appended to out; pieces are hashed on parallel threads.
void piece_digests(const unsigned char *buf, uint64_t n, std::vector<uint64_t> &out) {
    const uint64_t pieces = (n + kChunk - 1) / kChunk, base = out.size();
    out.resize(base + pieces);
    auto one = [&](uint64_t i) {
        const uint64_t b = i * kChunk;
        out[base + i] = xxh64(buf + b, (size_t)std::min(kChunk, n - b), 0);
    };
    if (pieces <= 1) {
        if (pieces) one(0);
        return;
    }
    std::vector<std::thread> th;
    for (uint64_t i = 1; i < pieces; ++i) th.emplace_back(one, i);
    one(0);
    for (auto &t : th) t.join();
}

uint64_t section_hash(const std::vector<uint64_t> &digests) {
    return xxh64(digests.data(), digests.size() * 8, 0);
}

uint64_t align_up(uint64_t x) { return (x + 4095u) & ~4095ull; }

struct Header {
    uint32_t n[3] = {0, 0, 0};
    float bound = 0.0f;
    std::string key;
    uint64_t off[4] = {0, 0, 0, 0}, bytes[4] = {0, 0, 0, 0}, hash[4] = {0, 0, 0, 0};
};

void section_sizes(Header &h) {
    h.bytes[0] = (uint64_t)h.n[0] * 16u;
    h.bytes[1] = (uint64_t)h.n[0] * 4096u * 16u;
    h.bytes[2] = (uint64_t)h.n[1] * 256u * 16u;
    h.bytes[3] = (uint64_t)h.n[2] * 256u;
    uint64_t o = kHeaderBytes;
    for (int k = 0; k < 4; ++k) {
        h.off[k] = o;
        o = align_up(o + h.bytes[k]);
    }
}

void encode(const Header &h, unsigned char *b) {
    std::memset(b, 0, kHeaderBytes);
    std::memcpy(b, kMagic, 8);
    const uint32_t w[8] = {kFormat, kHeaderBytes, h.n[0], h.n[1], h.n[2], 0u, (uint32_t)h.key.size(),
                           (uint32_t)kChunk};
    std::memcpy(b + 8, w, sizeof w);
    std::memcpy(b + 28, &h.bound, 4);
    std::memcpy(b + 40, h.off, 32);
    std::memcpy(b + 72, h.bytes, 32);
    std::memcpy(b + 104, h.hash, 32);
    std::memcpy(b + kKeyOff, h.key.data(), h.key.size());
    const uint64_t hh = xxh64(b, kHashOff, 0);
    std::memcpy(b + kHashOff, &hh, 8);
}

// WX_OK, WX_ERR_NOT_FOUND (no such file, or another key), WX_ERR_INVALID (not a
// layout file of this format, or a damaged header).
int read_header(FILE *f, const char *path, const std::string &key, const char *who, Header &h) {
    unsigned char b[kHeaderBytes];
    if (std::fread(b, 1, kHeaderBytes, f) != kHeaderBytes)
        return set_error(WX_ERR_INVALID, std::string(who) + ": " + path + " is shorter than a layout header");
    uint64_t hh;
    std::memcpy(&hh, b + kHashOff, 8);
    if (std::memcmp(b, kMagic, 8) != 0)
        return set_error(WX_ERR_INVALID, std::string(who) + ": " + path + " is not a grid-layout file");
    if (xxh64(b, kHashOff, 0) != hh)
        return set_error(WX_ERR_INVALID, std::string(who) + ": " + path + ": damaged header");
    uint32_t w[8];
    std::memcpy(w, b + 8, sizeof w);
    if (w[0] != kFormat || w[1] != kHeaderBytes || w[7] != (uint32_t)kChunk || w[6] > kKeyMax)
        return set_error(WX_ERR_INVALID, std::string(who) + ": " + path + ": unsupported layout format");
    h.n[0] = w[2], h.n[1] = w[3], h.n[2] = w[4];
    std::memcpy(&h.bound, b + 28, 4);
    h.key.assign(reinterpret_cast<const char *>(b + kKeyOff), w[6]);
    Header want = h;
    section_sizes(want);
    uint64_t off[4], bytes[4];
    std::memcpy(off, b + 40, 32);
    std::memcpy(bytes, b + 72, 32);
    std::memcpy(h.hash, b + 104, 32);
    for (int k = 0; k < 4; ++k)
        if (off[k] != want.off[k] || bytes[k] != want.bytes[k])
            return set_error(WX_ERR_INVALID, std::string(who) + ": " + path + ": section table does not match the counts");
    std::memcpy(h.off, off, 32);
    std::memcpy(h.bytes, bytes, 32);
    if (h.key != key)
        return set_error(WX_ERR_NOT_FOUND, std::string(who) + ": " + path + " holds the layout of another key");
    return WX_OK;
}

int check_key(const char *key, const char *who) {
    if (!key) return set_error(WX_ERR_INVALID, std::string(who) + ": NULL key");
    if (std::strlen(key) > kKeyMax)
        return set_error(WX_ERR_INVALID, std::string(who) + ": key longer than 2048 bytes");
    return WX_OK;
}

int open_for_read(const char *path, const char *who, FILE **f) {
    *f = std::fopen(path, "rb");
    if (*f) return WX_OK;
    const int e = errno;
    return set_error(e == ENOENT ? WX_ERR_NOT_FOUND : WX_ERR_INVALID,
                     std::string(who) + ": cannot open " + path + ": " + std::strerror(e));
}

// Pinned staging: two buffers, each with the event of the copy that last used it.
struct Staging {
    unsigned char *buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipStream_t st = nullptr;
    bool ok = false;
    Staging() {
        ok = hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess;
        for (int i = 0; i < 2 && ok; ++i)
            ok = hipHostMalloc(reinterpret_cast<void **>(&buf[i]), kStage, hipHostMallocDefault) == hipSuccess &&
                 hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) == hipSuccess;
    }
    ~Staging() {
        if (st) (void)hipStreamSynchronize(st);
        for (int i = 0; i < 2; ++i) {
            if (buf[i]) (void)hipHostFree(buf[i]);
            if (ev[i]) (void)hipEventDestroy(ev[i]);
        }
        if (st) (void)hipStreamDestroy(st);
    }
};

// Writes a layout whose sections come from host pointers (dev = false) or device
// pointers (dev = true) to path (through path + ".tmp", renamed when complete).
int write_layout(const char *path, const char *key, const uint32_t n[3], float bound, const void *const src[4],
                 bool dev, const char *who) {
    if (!path) return set_error(WX_ERR_INVALID, std::string(who) + ": NULL path");
    if (int rc = check_key(key, who)) return rc;
    Header h;
    std::memcpy(h.n, n, sizeof h.n);
    h.bound = bound;
    h.key = key;
    section_sizes(h);
    const std::string tmp = std::string(path) + ".tmp";
    FILE *f = std::fopen(tmp.c_str(), "wb");
    if (!f) return set_error(WX_ERR_INVALID, std::string(who) + ": cannot create " + tmp + ": " + std::strerror(errno));
    auto fail = [&](int code, const std::string &msg) {
        std::fclose(f);
        std::remove(tmp.c_str());
        return set_error(code, std::string(who) + ": " + msg);
    };
    std::vector<unsigned char> zeros(kHeaderBytes, 0);
    if (std::fwrite(zeros.data(), 1, kHeaderBytes, f) != kHeaderBytes) return fail(WX_ERR_INVALID, "write failed");
    Staging S;
    if (dev && !S.ok) return fail(WX_ERR_DEVICE, "pinned staging allocation failed");
    for (int k = 0; k < 4; ++k) {
        std::vector<uint64_t> dg;
        const uint64_t n_bytes = h.bytes[k];
        if (std::fseek(f, (long)h.off[k], SEEK_SET) != 0) return fail(WX_ERR_INVALID, "seek failed");
        if (!dev) {
            const unsigned char *p = static_cast<const unsigned char *>(src[k]);
            for (uint64_t b = 0; b < n_bytes; b += kStage) {
                const uint64_t m = std::min(kStage, n_bytes - b);
                piece_digests(p + b, m, dg);
                if (std::fwrite(p + b, 1, m, f) != m) return fail(WX_ERR_INVALID, "write failed");
            }
        } else {
            // copy of piece i + 1 in flight while piece i is hashed and written
            const unsigned char *p = static_cast<const unsigned char *>(src[k]);
            const uint64_t stages = (n_bytes + kStage - 1) / kStage;
            auto issue = [&](uint64_t i) {
                const uint64_t b = i * kStage, m = std::min(kStage, n_bytes - b);
                return hipMemcpyAsync(S.buf[i & 1], p + b, m, hipMemcpyDeviceToHost, S.st) == hipSuccess &&
                       hipEventRecord(S.ev[i & 1], S.st) == hipSuccess;
            };
            if (stages && !issue(0)) return fail(WX_ERR_DEVICE, "device-to-host copy failed");
            for (uint64_t i = 0; i < stages; ++i) {
                if (hipEventSynchronize(S.ev[i & 1]) != hipSuccess) return fail(WX_ERR_DEVICE, "copy failed");
                if (i + 1 < stages && !issue(i + 1)) return fail(WX_ERR_DEVICE, "device-to-host copy failed");
                const uint64_t m = std::min(kStage, n_bytes - i * kStage);
                piece_digests(S.buf[i & 1], m, dg);
                if (std::fwrite(S.buf[i & 1], 1, m, f) != m) return fail(WX_ERR_INVALID, "write failed");
            }
        }
        h.hash[k] = section_hash(dg);
    }
    // zeros up to the 4 KiB boundary after the last section (the file position is at
    // its end), then the header
    const uint64_t pad = align_up(h.off[3] + h.bytes[3]) - (h.off[3] + h.bytes[3]);
    if (pad && std::fwrite(zeros.data(), 1, pad, f) != pad) return fail(WX_ERR_INVALID, "write failed");
    unsigned char hb[kHeaderBytes];
    encode(h, hb);
    if (std::fseek(f, 0, SEEK_SET) != 0 || std::fwrite(hb, 1, kHeaderBytes, f) != kHeaderBytes)
        return fail(WX_ERR_INVALID, "write failed");
    if (std::fclose(f) != 0) {
        std::remove(tmp.c_str());
        return set_error(WX_ERR_INVALID, std::string(who) + ": close failed");
    }
    if (std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        return set_error(WX_ERR_INVALID, std::string(who) + ": cannot rename " + tmp + " to " + path);
    }
    return WX_OK;
}

}  // namespace

extern "C" {

int wx_layout_save(const wx_grid_desc *d, const char *path, const char *key) {
    if (!d) return set_error(WX_ERR_INVALID, "wx_layout_save: NULL descriptor");
    if ((d->n5 && (!d->root || !d->n5_records)) || (d->n4 && !d->n4_records) || (d->n3 && !d->leaves))
        return set_error(WX_ERR_INVALID, "wx_layout_save: NULL pool with a non-zero count");
    const uint32_t n[3] = {d->n5, d->n4, d->n3};
    const void *src[4] = {d->root, d->n5_records, d->n4_records, d->leaves};
    return write_layout(path, key, n, d->bound, src, false, "wx_layout_save");
}

int wx_grid_save(const wx_grid *g, const char *path, const char *key) {
    if (!g) return set_error(WX_ERR_INVALID, "wx_grid_save: NULL grid");
    const uint32_t n[3] = {g->n5, g->n4, g->n3};
    const void *src[4] = {g->mem[0], g->mem[1], g->mem[2], g->mem[3]};
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) return set_error(WX_ERR_DEVICE, "wx_grid_save: no HIP device");
    if (cur != g->device && hipSetDevice(g->device) != hipSuccess)
        return set_error(WX_ERR_DEVICE, "wx_grid_save: cannot select the grid's device");
    const int rc = write_layout(path, key, n, g->dev.bound, src, true, "wx_grid_save");
    if (cur != g->device) (void)hipSetDevice(cur);
    return rc;
}

int wx_layout_probe(const char *path, const char *key, uint32_t counts[3]) {
    if (!path || !counts) return set_error(WX_ERR_INVALID, "wx_layout_probe: NULL argument");
    if (int rc = check_key(key, "wx_layout_probe")) return rc;
    FILE *f = nullptr;
    if (int rc = open_for_read(path, "wx_layout_probe", &f)) return rc;
    Header h;
    int rc = read_header(f, path, key, "wx_layout_probe", h);
    if (rc == WX_OK) {
        // the file must reach the end of its last section
        if (std::fseek(f, 0, SEEK_END) != 0 || (uint64_t)std::ftell(f) < h.off[3] + h.bytes[3])
            rc = set_error(WX_ERR_INVALID, std::string("wx_layout_probe: ") + path + " is truncated");
        else
            std::memcpy(counts, h.n, sizeof h.n);
    }
    std::fclose(f);
    return rc;
}

int wx_grid_load(const char *path, const char *key, wx_grid **out) {
    const char *who = "wx_grid_load";
    if (!path || !out) return set_error(WX_ERR_INVALID, "wx_grid_load: NULL argument");
    if (int rc = check_key(key, who)) return rc;
    FILE *f = nullptr;
    if (int rc = open_for_read(path, who, &f)) return rc;
    struct Closer {
        FILE *f;
        ~Closer() { std::fclose(f); }
    } closer{f};
    Header h;
    if (int rc = read_header(f, path, key, who, h)) return rc;
    // the node sections on the host: checked by their hashes, then by grid_new's
    // bounds checks of every index the kernels derive from them
    std::vector<unsigned char> nodes[3];
    for (int k = 0; k < 3; ++k) {
        try {
            nodes[k].resize(h.bytes[k]);
        } catch (const std::bad_alloc &) {
            return set_error(WX_ERR_BUDGET, "wx_grid_load: host allocation failed");
        }
        if (std::fseek(f, (long)h.off[k], SEEK_SET) != 0 ||
            std::fread(nodes[k].data(), 1, h.bytes[k], f) != h.bytes[k])
            return set_error(WX_ERR_INVALID, std::string(who) + ": " + path + " is truncated");
        std::vector<uint64_t> dg;
        for (uint64_t b = 0; b < h.bytes[k]; b += kStage)
            piece_digests(nodes[k].data() + b, std::min(kStage, h.bytes[k] - b), dg);
        if (section_hash(dg) != h.hash[k])
            return set_error(WX_ERR_INVALID, std::string(who) + ": " + path + ": node section " + std::to_string(k) +
                                                 " does not match its hash");
    }
    wx_grid_desc d;
    d.n5 = h.n[0];
    d.n4 = h.n[1];
    d.n3 = h.n[2];
    d.bound = h.bound;
    d.root = reinterpret_cast<const int32_t *>(nodes[0].data());
    d.n5_records = reinterpret_cast<const uint32_t *>(nodes[1].data());
    d.n4_records = reinterpret_cast<const uint32_t *>(nodes[2].data());
    d.leaves = nullptr;
    wx_grid *g = nullptr;
    if (int rc = wx::grid_new(&d, who, &g)) return rc;
    // leaf words: read into pinned staging, hashed, copied up while the next piece is read
    Staging S;
    if (!S.ok) {
        wx_grid_destroy(g);
        return set_error(WX_ERR_DEVICE, "wx_grid_load: pinned staging allocation failed");
    }
    auto fail = [&](int code, const std::string &msg) {
        (void)hipStreamSynchronize(S.st);
        wx_grid_destroy(g);
        return set_error(code, std::string(who) + ": " + msg);
    };
    const uint64_t n_bytes = h.bytes[3], stages = (n_bytes + kStage - 1) / kStage;
    unsigned char *dst = static_cast<unsigned char *>(g->mem[3]);
    std::vector<uint64_t> dg;
    if (std::fseek(f, (long)h.off[3], SEEK_SET) != 0) return fail(WX_ERR_INVALID, "seek failed");
    for (uint64_t i = 0; i < stages; ++i) {
        const uint64_t b = i * kStage, m = std::min(kStage, n_bytes - b);
        if (i >= 2 && hipEventSynchronize(S.ev[i & 1]) != hipSuccess) return fail(WX_ERR_DEVICE, "copy failed");
        if (std::fread(S.buf[i & 1], 1, m, f) != m) return fail(WX_ERR_INVALID, std::string(path) + " is truncated");
        piece_digests(S.buf[i & 1], m, dg);
        if (hipMemcpyAsync(dst + b, S.buf[i & 1], m, hipMemcpyHostToDevice, S.st) != hipSuccess ||
            hipEventRecord(S.ev[i & 1], S.st) != hipSuccess)
            return fail(WX_ERR_DEVICE, "host-to-device copy failed");
    }
    if (hipStreamSynchronize(S.st) != hipSuccess) return fail(WX_ERR_DEVICE, "copy failed");
    if (section_hash(dg) != h.hash[3])
        return fail(WX_ERR_INVALID, std::string(path) + ": leaf section does not match its hash");
    *out = g;
    return WX_OK;
}

}  // extern "C"
