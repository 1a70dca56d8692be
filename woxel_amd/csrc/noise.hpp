// Seeded 3-octave value noise for the synthetic isosurface grids, shared by the
// host builder (vdb_host.cpp) and the device generator (noise_device.hip) so that
// both produce the same voxels bit for bit: every operation is an IEEE f32 add,
// multiply, divide or floor in a fixed order (both sides build with contraction off).
//
// The includer defines WX_NOISE_FN as the function qualifiers of its compiler
// (`inline` on the host, `__host__ __device__ inline` in HIP sources).
#pragma once

#include <cstdint>

#ifndef WX_NOISE_FN
#define WX_NOISE_FN inline
#endif

namespace wxnoise {

constexpr int kOctaves = 3;

// Hashed lattice value in [0, 1) per integer point.
WX_NOISE_FN float lattice(uint32_t seed, int32_t i, int32_t j, int32_t k) {
    uint32_t h = seed ^ 0x9E3779B9u;
    h ^= (uint32_t)i * 0x85EBCA6Bu;
    h = (h << 13) | (h >> 19);
    h ^= (uint32_t)j * 0xC2B2AE35u;
    h = (h << 13) | (h >> 19);
    h ^= (uint32_t)k * 0x27D4EB2Fu;
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    h ^= h >> 16;
    return (float)(h >> 8) * (1.0f / 16777216.0f);
}

WX_NOISE_FN float smooth(float t) { return t * t * (3.0f - 2.0f * t); }

// The lattice values at the corners of cell (i, j, k): l[b][a][d] at (i + d, j + a, k + b).
struct Cell {
    int32_t i, j, k;
    float l[2][2][2];
};

WX_NOISE_FN void cell_corners(uint32_t seed, int32_t i, int32_t j, int32_t k, Cell &q) {
    q.i = i, q.j = j, q.k = k;
    for (int b = 0; b < 2; ++b)
        for (int a = 0; a < 2; ++a) {
            q.l[b][a][0] = lattice(seed, i, j + a, k + b);
            q.l[b][a][1] = lattice(seed, i + 1, j + a, k + b);
        }
}

// Trilinear interpolation of a cell's corners with smoothstep weights.
WX_NOISE_FN float blend(const Cell &q, float u, float v, float w) {
    float c[2][2];
    for (int b = 0; b < 2; ++b)
        for (int a = 0; a < 2; ++a) {
            const float l0 = q.l[b][a][0], l1 = q.l[b][a][1];
            c[b][a] = l0 + (l1 - l0) * u;
        }
    const float e0 = c[0][0] + (c[0][1] - c[0][0]) * v, e1 = c[1][0] + (c[1][1] - c[1][0]) * v;
    return e0 + (e1 - e0) * w;
}

// Value noise at (x, y, z). With `cache`, the cell's corners are taken from it when it
// holds this cell and stored there otherwise: the same values, computed once per cell.
WX_NOISE_FN float vnoise(uint32_t seed, float x, float y, float z, Cell *cache = nullptr) {
    const float fx = __builtin_floorf(x), fy = __builtin_floorf(y), fz = __builtin_floorf(z);
    const int32_t i = (int32_t)fx, j = (int32_t)fy, k = (int32_t)fz;
    const float u = smooth(x - fx), v = smooth(y - fy), w = smooth(z - fz);
    if (cache) {
        if (cache->i != i || cache->j != j || cache->k != k) cell_corners(seed, i, j, k, *cache);
        return blend(*cache, u, v, w);
    }
    Cell q;
    cell_corners(seed, i, j, k, q);
    return blend(q, u, v, w);
}

// Three octaves of doubling frequency and halving amplitude, normalised to [0, 1);
// `cache` holds one cell per octave (see vnoise), or is NULL.
WX_NOISE_FN float noise3(uint32_t seed, float f0, float x, float y, float z, Cell *cache = nullptr) {
    float sum = 0.0f, amp = 1.0f, f = f0, norm = 0.0f;
    for (int o = 0; o < kOctaves; ++o) {
        sum += amp * vnoise(seed + 0x632BE5ABu * (uint32_t)o, x * f, y * f, z * f, cache ? cache + o : nullptr);
        norm += amp;
        amp *= 0.5f;
        f *= 2.0f;
    }
    return sum / norm;
}

// An empty cache: no lattice cell has these coordinates (i = INT32_MIN would need a
// coordinate below -2^31).
WX_NOISE_FN void clear_cache(Cell cache[kOctaves]) {
    for (int o = 0; o < kOctaves; ++o) cache[o].i = cache[o].j = cache[o].k = (int32_t)0x80000000u;
}

// |grad noise| <= 1.5 sqrt(3) per lattice unit per octave, and every voxel of a leaf
// lies within 3.5 sqrt(3) of the leaf centre (x + 3.5, y + 3.5, z + 3.5): a centre
// value farther than this from the threshold decides the whole leaf.
WX_NOISE_FN float leaf_bound(float f0) {
    float fsum = 0.0f, asum = 0.0f, amp = 1.0f, f = f0;
    for (int o = 0; o < kOctaves; ++o) {
        fsum += amp * f;
        asum += amp;
        amp *= 0.5f;
        f *= 2.0f;
    }
    return 1.5f * __builtin_sqrtf(3.0f) * 3.5f * __builtin_sqrtf(3.0f) * fsum / asum + 1e-4f;
}

// One isosurface: voxel g in [lo, hi) is active iff noise3(seed, f0, g) > thr.
struct Params {
    uint32_t seed;
    float f0, thr, bound;
    int32_t lo[3], hi[3];
};

// Leaf class from its centre value: 0 = no active voxel, 1 = every voxel active,
// 2 = decided voxel by voxel. b = the leaf's first voxel.
WX_NOISE_FN int leaf_class(const Params &p, const int32_t b[3], bool &inside_box) {
    inside_box = true;
    for (int a = 0; a < 3; ++a) inside_box &= b[a] >= p.lo[a] && b[a] + 8 <= p.hi[a];
    const float vc = noise3(p.seed, p.f0, (float)b[0] + 3.5f, (float)b[1] + 3.5f, (float)b[2] + 3.5f);
    if (inside_box && vc < p.thr - p.bound) return 0;
    if (inside_box && vc > p.thr + p.bound) return 1;
    return 2;
}

WX_NOISE_FN bool voxel_active(const Params &p, int32_t x, int32_t y, int32_t z, Cell *cache = nullptr) {
    if (x < p.lo[0] || x >= p.hi[0] || y < p.lo[1] || y >= p.hi[1] || z < p.lo[2] || z >= p.hi[2]) return false;
    return noise3(p.seed, p.f0, (float)x, (float)y, (float)z, cache) > p.thr;
}

}  // namespace wxnoise
