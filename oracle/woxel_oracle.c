/*
 * woxel_oracle.c — CPU oracle for the HDDA + grid-SDF traversal over VDB345.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product). See woxel_oracle.h
 * for the list of reference files restated here. Each function cites the
 * reference lines it follows. Built with -O2 -ffp-contract=off, no fast-math.
 */
#include "woxel_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* Node index math (data_structure.rs:17-92)                            */
/* ------------------------------------------------------------------ */

/* (c >> t) << t on i32 with an arithmetic right shift (data_structure.rs:73-75). */
static inline int32_t node_of(int32_t c, int t) {
    return (int32_t)((uint32_t)(c >> t) << t);
}

void or_global_to_node(const int32_t g[3], int total_log2, int32_t out[3]) {
    for (int i = 0; i < 3; ++i) out[i] = node_of(g[i], total_log2);
}

/* Node::global_to_offset (data_structure.rs:58-70). */
uint32_t or_global_to_offset(const int32_t g[3], int log2_d, int total_log2) {
    int child_total = total_log2 - log2_d;
    uint32_t m = (1u << total_log2) - 1u;
    uint32_t x = ((uint32_t)g[0] & m) >> child_total;
    uint32_t y = ((uint32_t)g[1] & m) >> child_total;
    uint32_t z = ((uint32_t)g[2] & m) >> child_total;
    return (x << (2 * log2_d)) | (y << log2_d) | z;
}

/* Node::offset_to_child (data_structure.rs:80-87). */
void or_offset_to_child(uint32_t off, int log2_d, uint32_t out[3]) {
    uint32_t dm = (1u << log2_d) - 1u;
    out[0] = off >> (2 * log2_d);
    out[1] = (off >> log2_d) & dm;
    out[2] = off & dm;
}

/* Node::child_to_offset (data_structure.rs:89-91). */
uint32_t or_child_to_offset(const uint32_t c[3], int log2_d) {
    return (c[0] << (2 * log2_d)) | (c[1] << log2_d) | c[2];
}

/* ------------------------------------------------------------------ */
/* Tree (data_structure.rs:94-269, vdb345.rs:8-12)                      */
/* ------------------------------------------------------------------ */

typedef struct OrN3 {
    uint64_t vmask[8]; /* value_mask */
    uint64_t isval[8]; /* data[o] is LeafData::Value (else LeafData::Tile) */
    uint64_t data[512];
} OrN3;

typedef struct OrN4 {
    uint64_t cmask[64], vmask[64];
    uint32_t tile[4096]; /* InternalData::Tile payload where child == NULL */
    OrN3 *child[4096];
} OrN4;

typedef struct OrN5 {
    uint64_t cmask[512], vmask[512];
    uint32_t tile[32768];
    OrN4 *child[32768];
} OrN5;

typedef struct OrRoot {
    int32_t key[3];
    OrN5 *node;
} OrRoot;

struct or_vdb {
    OrRoot *e;
    size_t n, cap;
};

or_vdb *or_vdb_new(void) {
    return (or_vdb *)calloc(1, sizeof(or_vdb));
}

void or_vdb_free(or_vdb *v) {
    if (!v) return;
    for (size_t r = 0; r < v->n; ++r) {
        OrN5 *n5 = v->e[r].node;
        for (int i = 0; i < 32768; ++i) {
            OrN4 *n4 = n5->child[i];
            if (!n4) continue;
            for (int j = 0; j < 4096; ++j) free(n4->child[j]);
            free(n4);
        }
        free(n5);
    }
    free(v->e);
    free(v);
}

static int key_eq(const int32_t a[3], const int32_t b[3]) {
    return a[0] == b[0] && a[1] == b[1] && a[2] == b[2];
}

/* Lexicographic order of [i32; 3] root keys (Rust's derived Ord). */
static int key_cmp(const void *pa, const void *pb) {
    const OrRoot *a = (const OrRoot *)pa, *b = (const OrRoot *)pb;
    for (int i = 0; i < 3; ++i) {
        if (a->key[i] < b->key[i]) return -1;
        if (a->key[i] > b->key[i]) return 1;
    }
    return 0;
}

static OrN5 *root_find(const or_vdb *v, const int32_t key[3]) {
    for (size_t r = 0; r < v->n; ++r)
        if (key_eq(v->e[r].key, key)) return v->e[r].node;
    return NULL;
}

/* Root entries in sorted key order (vdb345.rs:110 sorted_by_key). */
static OrRoot *root_sorted(const or_vdb *v) {
    OrRoot *s = (OrRoot *)malloc((v->n ? v->n : 1) * sizeof(OrRoot));
    memcpy(s, v->e, v->n * sizeof(OrRoot));
    qsort(s, v->n, sizeof(OrRoot), key_cmp);
    return s;
}

static inline void setbit(uint64_t *m, uint32_t b) { m[b >> 6] |= 1ull << (b & 63); }
static inline int getbit(const uint64_t *m, uint32_t b) { return (int)((m[b >> 6] >> (b & 63)) & 1ull); }

/* VDB345::set_voxel (vdb345.rs:26-66). New nodes hold Tile(0) everywhere. */
void or_vdb_set_voxel(or_vdb *v, int32_t x, int32_t y, int32_t z, uint32_t value_bits) {
    int32_t p[3] = {x, y, z}, key[3];
    or_global_to_node(p, 12, key);
    uint32_t b4 = or_global_to_offset(p, 5, 12);
    uint32_t b3 = or_global_to_offset(p, 4, 7);
    uint32_t b0 = or_global_to_offset(p, 3, 3);
    OrN5 *n5 = root_find(v, key);
    if (!n5) {
        if (v->n == v->cap) {
            v->cap = v->cap ? 2 * v->cap : 16;
            v->e = (OrRoot *)realloc(v->e, v->cap * sizeof(OrRoot));
        }
        n5 = (OrN5 *)calloc(1, sizeof(OrN5));
        memcpy(v->e[v->n].key, key, sizeof key);
        v->e[v->n].node = n5;
        v->n++;
    }
    if (!n5->child[b4]) n5->child[b4] = (OrN4 *)calloc(1, sizeof(OrN4));
    setbit(n5->cmask, b4);
    OrN4 *n4 = n5->child[b4];
    if (!n4->child[b3]) n4->child[b3] = (OrN3 *)calloc(1, sizeof(OrN3));
    setbit(n4->cmask, b3);
    OrN3 *n3 = n4->child[b3];
    setbit(n3->vmask, b0);
    setbit(n3->isval, b0);
    n3->data[b0] = value_bits;
}

/* VDB345::get_voxel (vdb345.rs:69-106). */
int or_vdb_get_voxel(const or_vdb *v, int32_t x, int32_t y, int32_t z, uint64_t *value, int *level) {
    int32_t p[3] = {x, y, z}, key[3];
    or_global_to_node(p, 12, key);
    *level = 0;
    OrN5 *n5 = root_find(v, key);
    if (!n5) {
        *value = 0; /* background */
        return OR_EP_BKGR;
    }
    uint32_t b4 = or_global_to_offset(p, 5, 12);
    OrN4 *n4 = n5->child[b4];
    if (!n4) {
        *value = n5->tile[b4];
        *level = 5;
        return OR_EP_INNR;
    }
    uint32_t b3 = or_global_to_offset(p, 4, 7);
    OrN3 *n3 = n4->child[b3];
    if (!n3) {
        *value = n4->tile[b3];
        *level = 4;
        return OR_EP_INNR;
    }
    uint32_t b0 = or_global_to_offset(p, 3, 3);
    *value = n3->data[b0];
    return getbit(n3->isval, b0) ? OR_EP_LEAF : OR_EP_OFFS;
}

/* VDB345::count_nodes (vdb345.rs:266-287). */
void or_vdb_count_nodes(const or_vdb *v, uint64_t out[3]) {
    out[0] = out[1] = out[2] = 0;
    for (size_t r = 0; r < v->n; ++r) {
        OrN5 *n5 = v->e[r].node;
        out[0]++;
        for (int i = 0; i < 32768; ++i) {
            OrN4 *n4 = n5->child[i];
            if (!n4) continue;
            out[1]++;
            for (int j = 0; j < 4096; ++j)
                if (n4->child[j]) out[2]++;
        }
    }
}

int or_vdb_leaf(const or_vdb *v, int32_t x, int32_t y, int32_t z, uint64_t data[512], uint64_t isval[8]) {
    int32_t p[3] = {x, y, z}, key[3];
    or_global_to_node(p, 12, key);
    OrN5 *n5 = root_find(v, key);
    if (!n5) return 0;
    OrN4 *n4 = n5->child[or_global_to_offset(p, 5, 12)];
    if (!n4) return 0;
    OrN3 *n3 = n4->child[or_global_to_offset(p, 4, 7)];
    if (!n3) return 0;
    memcpy(data, n3->data, sizeof n3->data);
    memcpy(isval, n3->isval, sizeof n3->isval);
    return 1;
}

void or_vdb_add_sphere(or_vdb *v, double cx, double cy, double cz, double r, const int32_t lo[3],
                       const int32_t hi[3], int value_mode, uint32_t value_bits) {
    double c[3] = {cx, cy, cz};
    int32_t a[3], b[3];
    for (int i = 0; i < 3; ++i) {
        double fa = floor(c[i] - r), fb = ceil(c[i] + r);
        a[i] = fa < (double)lo[i] ? lo[i] : (int32_t)fa;
        b[i] = fb > (double)(hi[i] - 1) ? hi[i] - 1 : (int32_t)fb;
    }
    double r2 = r * r;
    for (int32_t x = a[0]; x <= b[0]; ++x)
        for (int32_t y = a[1]; y <= b[1]; ++y)
            for (int32_t z = a[2]; z <= b[2]; ++z) {
                double dx = (double)x - cx, dy = (double)y - cy, dz = (double)z - cz;
                double d2 = dx * dx + dy * dy + dz * dz;
                if (d2 <= r2) {
                    uint32_t bits = value_bits;
                    if (value_mode == 1) {
                        float f = (float)(r - sqrt(d2));
                        memcpy(&bits, &f, 4);
                    }
                    or_vdb_set_voxel(v, x, y, z, bits);
                }
            }
}

void or_vdb_add_leaf_masks(or_vdb *v, const int32_t *origins, const uint64_t *masks, size_t n, uint32_t value_bits) {
    for (size_t i = 0; i < n; ++i)
        for (uint32_t o = 0; o < 512; ++o)
            if ((masks[8 * i + (o >> 6)] >> (o & 63)) & 1ull)
                or_vdb_set_voxel(v, origins[3 * i] + (int32_t)(o >> 6), origins[3 * i + 1] + (int32_t)((o >> 3) & 7),
                                 origins[3 * i + 2] + (int32_t)(o & 7), value_bits);
}

/* A node as read_tree_topology creates it (read.rs:303-340, new_from_header
 * data_structure.rs:142-153, 199-213): the node and its parents exist even
 * without children; level 5 / 4 value masks mark active tiles (tile payloads
 * stay Tile(0) until compute_sdf), a leaf's value mask marks its active voxels,
 * which read_tree_data turns into LeafData::Value (read.rs:559-570; value 0). */
void or_vdb_add_node(or_vdb *v, int level, const int32_t origin[3], const uint64_t *vmask) {
    int32_t key[3];
    or_global_to_node(origin, 12, key);
    OrN5 *n5 = root_find(v, key);
    if (!n5) {
        if (v->n == v->cap) {
            v->cap = v->cap ? 2 * v->cap : 16;
            v->e = (OrRoot *)realloc(v->e, v->cap * sizeof(OrRoot));
        }
        n5 = (OrN5 *)calloc(1, sizeof(OrN5));
        memcpy(v->e[v->n].key, key, sizeof key);
        v->e[v->n].node = n5;
        v->n++;
    }
    if (level == 5) {
        for (int w = 0; w < 512; ++w) n5->vmask[w] |= vmask[w];
        return;
    }
    uint32_t b4 = or_global_to_offset(origin, 5, 12);
    if (!n5->child[b4]) n5->child[b4] = (OrN4 *)calloc(1, sizeof(OrN4));
    setbit(n5->cmask, b4);
    OrN4 *n4 = n5->child[b4];
    if (level == 4) {
        for (int w = 0; w < 64; ++w) n4->vmask[w] |= vmask[w];
        return;
    }
    uint32_t b3 = or_global_to_offset(origin, 4, 7);
    if (!n4->child[b3]) n4->child[b3] = (OrN3 *)calloc(1, sizeof(OrN3));
    setbit(n4->cmask, b3);
    OrN3 *n3 = n4->child[b3];
    for (int w = 0; w < 8; ++w) {
        n3->vmask[w] |= vmask[w];
        n3->isval[w] |= vmask[w];
    }
}

/* ------------------------------------------------------------------ */
/* compute_sdf (vdb345.rs:290-628)                                      */
/* ------------------------------------------------------------------ */

/* Neighbour sets in the reference's push order (vdb345.rs:327-343). */
static int F_NB[13][3], B_NB[13][3];
static pthread_once_t nb_once = PTHREAD_ONCE_INIT;
static void nb_init(void) {
    int k = 0;
    for (int dy = -1; dy <= 1; ++dy)
        for (int dz = -1; dz <= 1; ++dz) {
            F_NB[k][0] = -1; F_NB[k][1] = dy; F_NB[k][2] = dz;
            B_NB[k][0] = 1;  B_NB[k][1] = dy; B_NB[k][2] = dz;
            ++k;
        }
    for (int dz = -1; dz <= 1; ++dz) {
        F_NB[k][0] = 0; F_NB[k][1] = -1; F_NB[k][2] = dz;
        B_NB[k][0] = 0; B_NB[k][1] = 1;  B_NB[k][2] = dz;
        ++k;
    }
    F_NB[k][0] = 0; F_NB[k][1] = 0; F_NB[k][2] = -1;
    B_NB[k][0] = 0; B_NB[k][1] = 0; B_NB[k][2] = 1;
}

static inline uint32_t min_u32(uint32_t a, uint32_t b) { return a < b ? a : b; }
static inline uint64_t min_u64(uint64_t a, uint64_t b) { return a < b ? a : b; }

/* One N5 tile (vdb345.rs:368-395 forward, :508-535 backward). */
static void sdf_n5_tile(const or_vdb *v, OrN5 *n5, uint32_t n4i, const int32_t origin5[3], int nb[13][3]) {
    uint32_t c5[3];
    or_offset_to_child(n4i, 5, c5);
    int32_t g[3], nn[3], ng[3];
    for (int i = 0; i < 3; ++i) g[i] = origin5[i] + (int32_t)c5[i] * 128;
    uint32_t *tv = &n5->tile[n4i];
    for (int k = 0; k < 13; ++k) {
        int32_t nglobal[3], nchild[3];
        for (int i = 0; i < 3; ++i) {
            nchild[i] = (int32_t)c5[i] + nb[k][i];
            nglobal[i] = g[i] + nb[k][i] * 128;
        }
        or_global_to_node(nglobal, 12, nn);
        or_global_to_node(g, 12, ng);
        if (key_eq(nn, ng)) {
            uint32_t nc[3] = {(uint32_t)nchild[0], (uint32_t)nchild[1], (uint32_t)nchild[2]};
            uint32_t nid = or_child_to_offset(nc, 5);
            if (n5->child[nid]) { *tv = 1; break; }
            *tv = min_u32(*tv, n5->tile[nid] + 1u);
            continue;
        }
        uint64_t val; int lvl;
        int ep = or_vdb_get_voxel(v, nglobal[0], nglobal[1], nglobal[2], &val, &lvl);
        *tv = (ep == OR_EP_INNR && lvl == 5) ? min_u32(*tv, (uint32_t)val + 1u) : 1u;
    }
}

/* One N4 tile (vdb345.rs:405-435 forward, :545-575 backward). */
static void sdf_n4_tile(const or_vdb *v, OrN4 *n4, uint32_t n3i, const int32_t g4[3], int nb[13][3]) {
    uint32_t c4[3];
    or_offset_to_child(n3i, 4, c4);
    int32_t g[3], nn[3], ng[3];
    for (int i = 0; i < 3; ++i) g[i] = g4[i] + (int32_t)c4[i] * 8;
    uint32_t *tv = &n4->tile[n3i];
    for (int k = 0; k < 13; ++k) {
        int32_t nglobal[3], nchild[3];
        for (int i = 0; i < 3; ++i) {
            nchild[i] = (int32_t)c4[i] + nb[k][i];
            nglobal[i] = g[i] + nb[k][i] * 8;
        }
        or_global_to_node(nglobal, 7, nn);
        or_global_to_node(g, 7, ng);
        if (key_eq(nn, ng)) {
            uint32_t nc[3] = {(uint32_t)nchild[0], (uint32_t)nchild[1], (uint32_t)nchild[2]};
            uint32_t nid = or_child_to_offset(nc, 4);
            if (n4->child[nid]) { *tv = 1; break; }
            *tv = min_u32(*tv, n4->tile[nid] + 1u);
            continue;
        }
        uint64_t val; int lvl;
        int ep = or_vdb_get_voxel(v, nglobal[0], nglobal[1], nglobal[2], &val, &lvl);
        *tv = (ep == OR_EP_INNR && lvl == 4) ? min_u32(*tv, (uint32_t)val + 1u) : 1u;
    }
}

/* One N3 voxel (vdb345.rs:444-477 forward, :586-619 backward). */
static void sdf_n3_voxel(const or_vdb *v, OrN3 *n3, uint32_t vi, const int32_t g3[3], int nb[13][3]) {
    if (getbit(n3->isval, vi)) return; /* LeafData::Value: skipped */
    uint32_t c3[3];
    or_offset_to_child(vi, 3, c3);
    int32_t g[3], nn[3], ng[3];
    for (int i = 0; i < 3; ++i) g[i] = g3[i] + (int32_t)c3[i];
    uint64_t *tv = &n3->data[vi];
    for (int k = 0; k < 13; ++k) {
        int32_t nglobal[3], nchild[3];
        for (int i = 0; i < 3; ++i) {
            nchild[i] = (int32_t)c3[i] + nb[k][i];
            nglobal[i] = g[i] + nb[k][i];
        }
        or_global_to_node(nglobal, 3, nn);
        or_global_to_node(g, 3, ng);
        if (key_eq(nn, ng)) {
            uint32_t nc[3] = {(uint32_t)nchild[0], (uint32_t)nchild[1], (uint32_t)nchild[2]};
            uint32_t nid = or_child_to_offset(nc, 3);
            if (getbit(n3->isval, nid)) { *tv = 1; break; }
            *tv = min_u64(*tv, n3->data[nid] + 1u);
            continue;
        }
        uint64_t val; int lvl;
        int ep = or_vdb_get_voxel(v, nglobal[0], nglobal[1], nglobal[2], &val, &lvl);
        *tv = (ep == OR_EP_OFFS) ? min_u64(*tv, val + 1u) : 1u;
    }
}

void or_vdb_compute_sdf(or_vdb *v) {
    pthread_once(&nb_once, nb_init);
    /* Initialise every tile and inactive voxel to MAX - 1 (vdb345.rs:291-325). */
    for (size_t r = 0; r < v->n; ++r) {
        OrN5 *n5 = v->e[r].node;
        for (int i = 0; i < 32768; ++i) {
            OrN4 *n4 = n5->child[i];
            if (!n4) { n5->tile[i] = UINT32_MAX - 1u; continue; }
            for (int j = 0; j < 4096; ++j) {
                OrN3 *n3 = n4->child[j];
                if (!n3) { n4->tile[j] = UINT32_MAX - 1u; continue; }
                for (int k = 0; k < 512; ++k)
                    if (!getbit(n3->isval, (uint32_t)k)) n3->data[k] = UINT64_MAX - 1u;
            }
        }
    }
    OrRoot *s = root_sorted(v);
    /* Forward pass (vdb345.rs:352-485): sorted keys, offsets ascending. */
    for (size_t r = 0; r < v->n; ++r) {
        OrN5 *n5 = s[r].node;
        const int32_t *o5 = s[r].key;
        for (uint32_t n4i = 0; n4i < 32768; ++n4i) {
            OrN4 *n4 = n5->child[n4i];
            if (!n4) { sdf_n5_tile(v, n5, n4i, o5, F_NB); continue; }
            uint32_t c5[3];
            or_offset_to_child(n4i, 5, c5);
            int32_t g4[3] = {o5[0] + (int32_t)c5[0] * 128, o5[1] + (int32_t)c5[1] * 128, o5[2] + (int32_t)c5[2] * 128};
            for (uint32_t n3i = 0; n3i < 4096; ++n3i) {
                OrN3 *n3 = n4->child[n3i];
                if (!n3) { sdf_n4_tile(v, n4, n3i, g4, F_NB); continue; }
                uint32_t c4[3];
                or_offset_to_child(n3i, 4, c4);
                int32_t g3[3] = {g4[0] + (int32_t)c4[0] * 8, g4[1] + (int32_t)c4[1] * 8, g4[2] + (int32_t)c4[2] * 8};
                for (uint32_t vi = 0; vi < 512; ++vi) sdf_n3_voxel(v, n3, vi, g3, F_NB);
            }
        }
    }
    /* Backward pass (vdb345.rs:487-627): exactly the reverse order. */
    for (size_t rr = v->n; rr-- > 0;) {
        OrN5 *n5 = s[rr].node;
        const int32_t *o5 = s[rr].key;
        for (uint32_t n4i = 32768; n4i-- > 0;) {
            OrN4 *n4 = n5->child[n4i];
            if (!n4) { sdf_n5_tile(v, n5, n4i, o5, B_NB); continue; }
            uint32_t c5[3];
            or_offset_to_child(n4i, 5, c5);
            int32_t g4[3] = {o5[0] + (int32_t)c5[0] * 128, o5[1] + (int32_t)c5[1] * 128, o5[2] + (int32_t)c5[2] * 128};
            for (uint32_t n3i = 4096; n3i-- > 0;) {
                OrN3 *n3 = n4->child[n3i];
                if (!n3) { sdf_n4_tile(v, n4, n3i, g4, B_NB); continue; }
                uint32_t c4[3];
                or_offset_to_child(n3i, 4, c4);
                int32_t g3[3] = {g4[0] + (int32_t)c4[0] * 8, g4[1] + (int32_t)c4[1] * 8, g4[2] + (int32_t)c4[2] * 8};
                for (uint32_t vi = 512; vi-- > 0;) sdf_n3_voxel(v, n3, vi, g3, B_NB);
            }
        }
    }
    free(s);
}

/* ------------------------------------------------------------------ */
/* Reference layout: origins / masks / atlas (vdb345.rs:108-264)        */
/* ------------------------------------------------------------------ */

/* arr32_from_arr64 (vdb345.rs:673-682): u64 word -> (lo, hi) u32 pair. */
static void arr32_from_arr64(const uint64_t *a, size_t n64, uint32_t *out) {
    for (size_t i = 0; i < n64; ++i) {
        out[2 * i] = (uint32_t)a[i];
        out[2 * i + 1] = (uint32_t)(a[i] >> 32);
    }
}

or_layout *or_layout_build(const or_vdb *v) {
    uint64_t cnt[3];
    or_vdb_count_nodes(v, cnt);
    or_layout *l = (or_layout *)calloc(1, sizeof(or_layout));
    l->n5 = (uint32_t)cnt[0];
    l->n4 = (uint32_t)cnt[1];
    l->n3 = (uint32_t)cnt[2];
    size_t a5 = cnt[0] ? cnt[0] : 1, a4 = cnt[1] ? cnt[1] : 1, a3 = cnt[2] ? cnt[2] : 1;
    l->origins = (int32_t *)calloc(a5 * 4, sizeof(int32_t));
    l->kids5 = (uint32_t *)calloc(a5 * 1024, 4);
    l->vals5 = (uint32_t *)calloc(a5 * 1024, 4);
    l->kids4 = (uint32_t *)calloc(a4 * 128, 4);
    l->vals4 = (uint32_t *)calloc(a4 * 128, 4);
    l->vals3 = (uint32_t *)calloc(a3 * 16, 4);
    l->atlas5 = (uint32_t *)calloc(a5 * 32768, 4);
    l->atlas4 = (uint32_t *)calloc(a4 * 4096, 4);
    l->atlas3 = (uint32_t *)calloc(a3 * 512, 4);
    OrRoot *s = root_sorted(v);
    uint32_t i5 = 0, i4 = 0, i3 = 0;
    for (size_t r = 0; r < v->n; ++r, ++i5) {
        OrN5 *n5 = s[r].node;
        memcpy(&l->origins[4 * i5], s[r].key, 3 * sizeof(int32_t));
        for (uint32_t o5 = 0; o5 < 32768; ++o5) {
            OrN4 *n4 = n5->child[o5];
            if (!n4) { l->atlas5[(size_t)i5 * 32768 + o5] = n5->tile[o5]; continue; }
            for (uint32_t o4 = 0; o4 < 4096; ++o4) {
                OrN3 *n3 = n4->child[o4];
                if (!n3) { l->atlas4[(size_t)i4 * 4096 + o4] = n4->tile[o4]; continue; }
                for (uint32_t o3 = 0; o3 < 512; ++o3)
                    l->atlas3[(size_t)i3 * 512 + o3] = (uint32_t)n3->data[o3]; /* value bits or (offset as u32) */
                arr32_from_arr64(n3->vmask, 8, &l->vals3[(size_t)i3 * 16]);
                l->atlas4[(size_t)i4 * 4096 + o4] = i3;
                i3++;
            }
            arr32_from_arr64(n4->vmask, 64, &l->vals4[(size_t)i4 * 128]);
            arr32_from_arr64(n4->cmask, 64, &l->kids4[(size_t)i4 * 128]);
            l->atlas5[(size_t)i5 * 32768 + o5] = i4;
            i4++;
        }
        arr32_from_arr64(n5->vmask, 512, &l->vals5[(size_t)i5 * 1024]);
        arr32_from_arr64(n5->cmask, 512, &l->kids5[(size_t)i5 * 1024]);
    }
    free(s);
    return l;
}

static void *dup_words(const void *src, size_t n, size_t elem) {
    void *d = calloc(n ? n : 1, elem);
    if (n) memcpy(d, src, n * elem);
    return d;
}

or_layout *or_layout_from_arrays(uint32_t n5, uint32_t n4, uint32_t n3, const int32_t *origins,
                                 const uint32_t *kids5, const uint32_t *vals5, const uint32_t *kids4,
                                 const uint32_t *vals4, const uint32_t *vals3, const uint32_t *atlas5,
                                 const uint32_t *atlas4, const uint32_t *atlas3) {
    or_layout *l = (or_layout *)calloc(1, sizeof(or_layout));
    l->n5 = n5; l->n4 = n4; l->n3 = n3;
    l->origins = (int32_t *)dup_words(origins, (size_t)n5 * 4, 4);
    l->kids5 = (uint32_t *)dup_words(kids5, (size_t)n5 * 1024, 4);
    l->vals5 = (uint32_t *)dup_words(vals5, (size_t)n5 * 1024, 4);
    l->kids4 = (uint32_t *)dup_words(kids4, (size_t)n4 * 128, 4);
    l->vals4 = (uint32_t *)dup_words(vals4, (size_t)n4 * 128, 4);
    l->vals3 = (uint32_t *)dup_words(vals3, (size_t)n3 * 16, 4);
    l->atlas5 = (uint32_t *)dup_words(atlas5, (size_t)n5 * 32768, 4);
    l->atlas4 = (uint32_t *)dup_words(atlas4, (size_t)n4 * 4096, 4);
    l->atlas3 = (uint32_t *)dup_words(atlas3, (size_t)n3 * 512, 4);
    return l;
}

void or_layout_free(or_layout *l) {
    if (!l) return;
    free(l->origins); free(l->kids5); free(l->vals5); free(l->kids4); free(l->vals4);
    free(l->vals3); free(l->atlas5); free(l->atlas4); free(l->atlas3);
    free(l);
}

/* ------------------------------------------------------------------ */
/* Traversal (raycast.comp.wgsl:70-126, 344-519)                        */
/* ------------------------------------------------------------------ */

/* WGSL min/max as selects: min(a, b) = b < a ? b : a; max(a, b) = a < b ? b : a. */
static inline float smin(float a, float b) { return b < a ? b : a; }
static inline float smax(float a, float b) { return a < b ? b : a; }

/* f32 -> i32 conversion: truncation with saturation, NaN -> 0 (the gfx9
 * v_cvt_i32_f32 rule; WGSL clamps out-of-range values). */
static inline int32_t f2i(float x) {
    if (x != x) return 0;
    if (x >= 2147483648.0f) return INT32_MAX;
    if (x <= -2147483648.0f) return INT32_MIN;
    return (int32_t)x;
}

typedef struct Parent { int32_t origin[3]; uint32_t idx; } Parent;

/* VdbLeaf (raycast.comp.wgsl:349-354) plus lookup counters. */
typedef struct Leaf {
    uint32_t dist, num_parents;
    Parent parents[3];
    uint32_t l5, l4, l3;
} Leaf;

static inline int org_eq(const int32_t a[3], const int32_t b[3]) { return key_eq(a, b); }

static void from_node3(const or_layout *L, const int32_t pos[3], Leaf *lf);
static void from_node4(const or_layout *L, const int32_t pos[3], Leaf *lf);
static void from_node5(const or_layout *L, const int32_t pos[3], Leaf *lf);

/* get_vdb_leaf_from_nothing (raycast.comp.wgsl:398-413): linear scan of origins. */
static void from_nothing(const or_layout *L, const int32_t pos[3], Leaf *lf) {
    int32_t g5[3];
    or_global_to_node(pos, 12, g5);
    for (uint32_t i = 0; i < L->n5; ++i) {
        if (org_eq(g5, &L->origins[4 * i])) {
            memcpy(lf->parents[0].origin, g5, sizeof g5);
            lf->parents[0].idx = i;
            lf->num_parents = 1;
            from_node5(L, pos, lf);
            return;
        }
    }
    lf->dist = 1;
    lf->num_parents = 0;
}

/* get_vdb_leaf_from_node5 (raycast.comp.wgsl:415-444). */
static void from_node5(const or_layout *L, const int32_t pos[3], Leaf *lf) {
    lf->l5++;
    uint32_t loc[3], ch[3];
    for (int i = 0; i < 3; ++i) { loc[i] = (uint32_t)pos[i] & 4095u; ch[i] = loc[i] >> 7; }
    uint32_t off = or_child_to_offset(ch, 5);
    uint32_t idx = lf->parents[0].idx;
    int in_kid = (L->kids5[(size_t)idx * 1024 + (off >> 5)] >> (off & 31)) & 1u;
    int in_val = (L->vals5[(size_t)idx * 1024 + (off >> 5)] >> (off & 31)) & 1u;
    uint32_t node4_idx = L->atlas5[(size_t)idx * 32768 + off];
    if (in_val) { lf->dist = 0; lf->num_parents = 1; return; }
    if (!in_kid) { lf->dist = node4_idx; lf->num_parents = 1; return; }
    or_global_to_node(pos, 7, lf->parents[1].origin);
    lf->parents[1].idx = node4_idx;
    lf->num_parents = 2;
    from_node4(L, pos, lf);
}

/* get_vdb_leaf_from_node4 (raycast.comp.wgsl:446-475). */
static void from_node4(const or_layout *L, const int32_t pos[3], Leaf *lf) {
    lf->l4++;
    uint32_t loc[3], ch[3];
    for (int i = 0; i < 3; ++i) { loc[i] = (uint32_t)pos[i] & 127u; ch[i] = loc[i] >> 3; }
    uint32_t off = or_child_to_offset(ch, 4);
    uint32_t idx = lf->parents[1].idx;
    int in_kid = (L->kids4[(size_t)idx * 128 + (off >> 5)] >> (off & 31)) & 1u;
    int in_val = (L->vals4[(size_t)idx * 128 + (off >> 5)] >> (off & 31)) & 1u;
    uint32_t node3_idx = L->atlas4[(size_t)idx * 4096 + off];
    if (in_val) { lf->dist = 0; lf->num_parents = 2; return; }
    if (!in_kid) { lf->dist = node3_idx; lf->num_parents = 2; return; }
    or_global_to_node(pos, 3, lf->parents[2].origin);
    lf->parents[2].idx = node3_idx;
    lf->num_parents = 3;
    from_node3(L, pos, lf);
}

/* get_vdb_leaf_from_node3 (raycast.comp.wgsl:477-494). */
static void from_node3(const or_layout *L, const int32_t pos[3], Leaf *lf) {
    lf->l3++;
    uint32_t loc[3];
    for (int i = 0; i < 3; ++i) loc[i] = (uint32_t)pos[i] & 7u;
    uint32_t off = or_child_to_offset(loc, 3);
    uint32_t idx = lf->parents[2].idx;
    int in_val = (L->vals3[(size_t)idx * 16 + (off >> 5)] >> (off & 31)) & 1u;
    uint32_t voxel = L->atlas3[(size_t)idx * 512 + off];
    if (in_val) { lf->dist = 0; lf->num_parents = 3; return; }
    lf->dist = voxel;
    lf->num_parents = 3;
}

/* get_vdb_leaf_from_leaf (raycast.comp.wgsl:360-396): reuse the deepest cached parent. */
static void from_leaf(const or_layout *L, const int32_t pos[3], Leaf *lf) {
    int32_t g[3];
    if (lf->num_parents == 3) {
        or_global_to_node(pos, 3, g);
        if (org_eq(lf->parents[2].origin, g)) { from_node3(L, pos, lf); return; }
        or_global_to_node(pos, 7, g);
        if (org_eq(lf->parents[1].origin, g)) { from_node4(L, pos, lf); return; }
        or_global_to_node(pos, 12, g);
        if (org_eq(lf->parents[0].origin, g)) { from_node5(L, pos, lf); return; }
        from_nothing(L, pos, lf);
        return;
    }
    if (lf->num_parents == 2) {
        or_global_to_node(pos, 7, g);
        if (org_eq(lf->parents[1].origin, g)) { from_node4(L, pos, lf); return; }
        or_global_to_node(pos, 12, g);
        if (org_eq(lf->parents[0].origin, g)) { from_node5(L, pos, lf); return; }
        from_nothing(L, pos, lf);
        return;
    }
    if (lf->num_parents == 1) {
        or_global_to_node(pos, 12, g);
        if (org_eq(lf->parents[0].origin, g)) { from_node5(L, pos, lf); return; }
        from_nothing(L, pos, lf);
        return;
    }
    from_nothing(L, pos, lf);
}

static const float SCALE[4] = {1.f, 8.f, 128.f, 4096.f}; /* raycast.comp.wgsl:83 */

static void fill_hit(or_hit *hit, or_diag *diag, uint32_t state, const float p[3], const int mask[3],
                     const float step[3], const float src[3], const float dir[3], const uint32_t steps[4],
                     const Leaf *lf, uint32_t flags) {
    for (int c = 0; c < 3; ++c) {
        hit->voxel[c] = f2i(floorf(p[c]));
        hit->normal[c] = mask[c] ? (int8_t)(-(int)step[c]) : (int8_t)0;
    }
    float t = (p[0] - src[0]) * dir[0];
    t = t + (p[1] - src[1]) * dir[1];
    t = t + (p[2] - src[2]) * dir[2];
    hit->t = t;
    for (int k = 0; k < 4; ++k) hit->steps[k] = (uint16_t)steps[k];
    hit->state = (uint8_t)state;
    hit->lookups = lf->l5 | (lf->l4 << 10) | (lf->l3 << 20);
    if (diag) {
        memcpy(diag->p, p, sizeof diag->p);
        diag->flags = flags;
    }
}

/* Entry clip (extension; identical in the product kernel): a ray whose origin
 * lies outside [-bound, bound]^3 starts at its slab entry point, clamped into
 * the box. Returns 0 when the ray misses the box. */
static int entry_clip(const float src[3], const float dir[3], const float idir[3], float bound, float out[3]) {
    float tn = -INFINITY, tf = INFINITY;
    for (int c = 0; c < 3; ++c) {
        float a = (-bound - src[c]) * idir[c];
        float b = (bound - src[c]) * idir[c];
        tn = smax(tn, smin(a, b));
        tf = smin(tf, smax(a, b));
    }
    if (!(tn <= tf) || tf < 0.0f) return 0;
    for (int c = 0; c < 3; ++c) {
        float q = src[c] + tn * dir[c];
        out[c] = smin(smax(q, -bound), bound);
    }
    return 1;
}

/* hdda_ray (raycast.comp.wgsl:84-126). */
void or_hdda_ray(const or_layout *L, const float src[3], const float dir[3], float bound, int clip,
                 or_hit *hit, or_diag *diag) {
    float p[3], step[3], step01[3], idir[3];
    int mask[3] = {0, 0, 0};
    uint32_t steps[4] = {0, 0, 0, 0};
    uint32_t flags = 0;
    Leaf lf;
    memset(&lf, 0, sizeof lf);
    for (int c = 0; c < 3; ++c) {
        p[c] = src[c];
        step[c] = dir[c] < 0.0f ? -1.0f : 1.0f; /* sign11 (:70-76) */
        step01[c] = smax(0.0f, step[c]);
        idir[c] = 1.0f / dir[c];
        if (dir[c] == 0.0f) flags |= OR_F_ZERODIR;
    }
    if (clip && (bound < fabsf(src[0]) || bound < fabsf(src[1]) || bound < fabsf(src[2]))) {
        flags |= OR_F_CLIPPED;
        if (!entry_clip(src, dir, idir, bound, p)) {
            flags |= OR_F_CLIPMISS;
            fill_hit(hit, diag, OR_OOB, src, mask, step, src, dir, steps, &lf, flags);
            return;
        }
    }
    for (uint32_t i = 0; i < 1000u; ++i) {
        int32_t pos[3] = {f2i(floorf(p[0])), f2i(floorf(p[1])), f2i(floorf(p[2]))};
        from_leaf(L, pos, &lf);
        steps[lf.num_parents]++;
        if (lf.dist == 0u) {
            fill_hit(hit, diag, OR_HIT, p, mask, step, src, dir, steps, &lf, flags);
            return;
        }
        if (bound < fabsf(p[0]) || bound < fabsf(p[1]) || bound < fabsf(p[2])) {
            fill_hit(hit, diag, OR_OOB, p, mask, step, src, dir, steps, &lf, flags);
            return;
        }
        float size = (float)lf.dist;
        size = size * SCALE[3 - lf.num_parents];
        float tmax[3];
        for (int c = 0; c < 3; ++c) {
            float m = p[c] - size * floorf(p[c] / size); /* modulo_vec3f (:78-80) */
            tmax[c] = idir[c] * (size * step01[c] - m);
        }
        float tmin = smin(smin(tmax[0], tmax[1]), tmax[2]);
        for (int c = 0; c < 3; ++c) p[c] = p[c] + tmin * dir[c];
        mask[0] = (tmax[0] <= tmax[1]) & (tmax[0] <= tmax[2]);
        mask[1] = (tmax[1] <= tmax[2]) & (tmax[1] <= tmax[0]);
        mask[2] = (tmax[2] <= tmax[0]) & (tmax[2] <= tmax[1]);
        for (int c = 0; c < 3; ++c) p[c] = p[c] + (4e-4f * step[c]) * (float)mask[c];
        if (mask[0] + mask[1] + mask[2] > 1) flags |= OR_F_TIE;
        if (!isfinite(tmin) || !isfinite(p[0]) || !isfinite(p[1]) || !isfinite(p[2])) flags |= OR_F_NONFINITE;
    }
    fill_hit(hit, diag, OR_MAXSTEPS, p, mask, step, src, dir, steps, &lf, flags);
}

/* cp_main (raycast.comp.wgsl:62-63): p = gid + offset; dir = normalize(p.x*u + p.y*mv + wp). */
void or_camera_ray(const float cam[16], uint32_t px, uint32_t py, float offset, float dir[3]) {
    const float *u = cam + 4, *mv = cam + 8, *wp = cam + 12;
    float fx = (float)px + offset, fy = (float)py + offset;
    float v[3];
    for (int c = 0; c < 3; ++c) v[c] = (fx * u[c] + fy * mv[c]) + wp[c];
    float len = sqrtf((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
    for (int c = 0; c < 3; ++c) dir[c] = v[c] / len;
}

/* ComputeState::build (compute_state.rs:87-131) with look_at_rh (camera.rs:31-35).
 * The camera-to-world columns of a look-at view are (s, u, -f), eye; they are
 * taken directly rather than through a numeric 4x4 inverse. */
void or_camera_basis(const float eye[3], const float target[3], const float up[3], float aspect,
                     float fovy_deg, float width, float cam[16]) {
    float f[3], s[3], u[3];
    for (int c = 0; c < 3; ++c) f[c] = target[c] - eye[c];
    float fl = sqrtf((f[0] * f[0] + f[1] * f[1]) + f[2] * f[2]);
    float fi = 1.0f / fl;
    for (int c = 0; c < 3; ++c) f[c] = f[c] * fi;
    s[0] = f[1] * up[2] - f[2] * up[1];
    s[1] = f[2] * up[0] - f[0] * up[2];
    s[2] = f[0] * up[1] - f[1] * up[0];
    float sl = sqrtf((s[0] * s[0] + s[1] * s[1]) + s[2] * s[2]);
    float si = 1.0f / sl;
    for (int c = 0; c < 3; ++c) s[c] = s[c] * si;
    u[0] = s[1] * f[2] - s[2] * f[1];
    u[1] = s[2] * f[0] - s[0] * f[2];
    u[2] = s[0] * f[1] - s[1] * f[0];
    /* c2w columns: x = s, y = u, z = w = -f. */
    float height = width / aspect;
    float half_w = -width / 2.0f, half_h = height / 2.0f;
    float rad = fovy_deg * (3.14159265358979323846f / 180.0f); /* f32::to_radians */
    float tn = tanf(rad * 0.5f);
    for (int c = 0; c < 3; ++c) {
        float w = -f[c];
        cam[c] = eye[c];
        cam[4 + c] = s[c];
        cam[8 + c] = -u[c];
        /* wp = (-W/2) u + (H/2) v - w (H/2) / tan(fovy/2), component-wise. */
        cam[12 + c] = (half_w * s[c] + half_h * u[c]) - (w * half_h) / tn;
    }
    cam[3] = 0.0f; cam[7] = 0.0f; cam[11] = 0.0f; cam[15] = 0.0f;
}

/* Shadow ray (raycast.comp.wgsl:188-190), point light instead of a sun direction. */
void or_shadow_ray(const float p[3], const float dir[3], const int8_t normal[3], const float light[3],
                   float o_out[3], float d_out[3]) {
    float v[3];
    for (int c = 0; c < 3; ++c) {
        /* normal = -step*mask, so p - (4e-2*step)*mask = p + 4e-2*normal when mask is set. */
        float step = dir[c] < 0.0f ? -1.0f : 1.0f;
        float m = normal[c] != 0 ? 1.0f : 0.0f;
        o_out[c] = p[c] - (4e-2f * step) * m;
    }
    for (int c = 0; c < 3; ++c) v[c] = light[c] - o_out[c];
    float len = sqrtf((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
    for (int c = 0; c < 3; ++c) d_out[c] = v[c] / len;
}

/* ------------------------------------------------------------------ */
/* Threaded drivers                                                     */
/* ------------------------------------------------------------------ */

typedef struct Job {
    const or_layout *L;
    const float *cam;
    const float *rays;
    uint32_t w, h;
    size_t n;
    float offset, bound;
    int clip;
    or_hit *out;
    or_diag *diag;
    size_t next;
    pthread_mutex_t mu;
} Job;

static size_t job_grab(Job *j, size_t chunk, size_t *end) {
    pthread_mutex_lock(&j->mu);
    size_t b = j->next;
    j->next = b + chunk < j->n ? b + chunk : j->n;
    *end = j->next;
    pthread_mutex_unlock(&j->mu);
    return b;
}

static void *cam_worker(void *arg) {
    Job *j = (Job *)arg;
    for (;;) {
        size_t e, b = job_grab(j, 1024, &e);
        if (b >= e) break;
        for (size_t k = b; k < e; ++k) {
            uint32_t px = (uint32_t)(k % j->w), py = (uint32_t)(k / j->w);
            float d[3];
            or_camera_ray(j->cam, px, py, j->offset, d);
            or_hdda_ray(j->L, j->cam, d, j->bound, j->clip, &j->out[k], j->diag ? &j->diag[k] : NULL);
        }
    }
    return NULL;
}

static void *ray_worker(void *arg) {
    Job *j = (Job *)arg;
    for (;;) {
        size_t e, b = job_grab(j, 1024, &e);
        if (b >= e) break;
        for (size_t k = b; k < e; ++k)
            or_hdda_ray(j->L, &j->rays[6 * k], &j->rays[6 * k + 3], j->bound, j->clip, &j->out[k],
                        j->diag ? &j->diag[k] : NULL);
    }
    return NULL;
}

static void run_job(Job *j, void *(*fn)(void *), int nthreads) {
    if (nthreads < 1) nthreads = 1;
    pthread_mutex_init(&j->mu, NULL);
    pthread_t *t = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
    for (int i = 0; i < nthreads; ++i) pthread_create(&t[i], NULL, fn, j);
    for (int i = 0; i < nthreads; ++i) pthread_join(t[i], NULL);
    free(t);
    pthread_mutex_destroy(&j->mu);
}

/* ------------------------------------------------------------------ */
/* Shading: ray_trace, reflect_ray2, reflect_ray1 (raycast.comp.wgsl:144-342) */
/* ------------------------------------------------------------------ */

/* WGSL vector arithmetic in the order the shader writes it; dot = (x+y)+z,
 * normalize = v / sqrt(dot(v,v)), mix(a,b,t) = a*(1-t) + b*t, max(a,b) = a<b?b:a,
 * min(a,b) = b<a?b:a, f32 remainder = fmodf. */
typedef struct { float x, y, z; } V3;
static V3 V(float x, float y, float z) { V3 r = {x, y, z}; return r; }
static V3 vadd(V3 a, V3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static V3 vsub(V3 a, V3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static V3 vmul(V3 a, V3 b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
static V3 vsc(V3 a, float s) { return V(a.x * s, a.y * s, a.z * s); }
static V3 sv(float s, V3 a) { return V(s * a.x, s * a.y, s * a.z); }
static V3 vadds(V3 a, float s) { return V(a.x + s, a.y + s, a.z + s); }
static float vdot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static V3 vnorm(V3 a) { float l = sqrtf(vdot(a, a)); return V(a.x / l, a.y / l, a.z / l); }
static V3 vmix(V3 a, V3 b, float t) {
    float u = 1.0f - t;
    return V(a.x * u + b.x * t, a.y * u + b.y * t, a.z * u + b.z * t);
}
static V3 vneg(V3 a) { return V(-a.x, -a.y, -a.z); }

static const float K_D = 0.7f, K_A = 0.3f, REFLECTIVITY = 0.9f, WALL_I = 0.1f;
#define BASE_COLOR V(0.4f, 0.2f, 0.2f)
#define AMBIENT_COLOR V(0.4f, 0.4f, 0.3f)

typedef struct {
    const or_layout *L;
    float bound;
    int clip;
    const or_shade *s;
} Shade;

/* HDDAout of one hdda_ray call. */
typedef struct {
    int state;
    V3 p, mask;
    float i;
} Hdda;

static Hdda hdda(const Shade *c, V3 src, V3 dir) {
    or_hit h;
    or_diag dg;
    float o[3] = {src.x, src.y, src.z}, d[3] = {dir.x, dir.y, dir.z};
    or_hdda_ray(c->L, o, d, c->bound, c->clip, &h, &dg);
    Hdda r;
    r.state = h.state;
    r.p = V(dg.p[0], dg.p[1], dg.p[2]);
    r.mask = V(h.normal[0] != 0 ? 1.0f : 0.0f, h.normal[1] != 0 ? 1.0f : 0.0f, h.normal[2] != 0 ? 1.0f : 0.0f);
    uint32_t n = (uint32_t)h.steps[0] + h.steps[1] + h.steps[2] + h.steps[3];
    r.i = (float)(n ? n - 1u : 0u); /* the iteration of return; 0 for a ray the entry clip discards */
    return r;
}

static V3 sign11v(V3 d) { return V(d.x < 0.0f ? -1.0f : 1.0f, d.y < 0.0f ? -1.0f : 1.0f, d.z < 0.0f ? -1.0f : 1.0f); }
static V3 sun_neg(const Shade *c) { return vneg(V(c->s->sun_dir[0], c->s->sun_dir[1], c->s->sun_dir[2])); }
static V3 sun_rgb(const Shade *c) { return V(c->s->sun_color[0], c->s->sun_color[1], c->s->sun_color[2]); }

/* hit.p - 4e-2 * step * vec3<f32>(hit.mask) */
static V3 offset_origin(Hdda h, V3 step) { return vsub(h.p, vmul(sv(4e-2f, step), h.mask)); }

/* the walls of reflect_ray2/1 (:262-274, :322-334) */
static V3 walls(V3 N) {
    V3 Np = V(smax(0.0f, N.x), smax(0.0f, N.y), smax(0.0f, N.z));
    V3 Nn = vneg(V(smin(0.0f, N.x), smin(0.0f, N.y), smin(0.0f, N.z)));
    V3 c = vsc(V(WALL_I, 0.0f, 0.0f), Np.x);
    c = vadd(c, vsc(V(0.0f, WALL_I, 0.0f), Np.y));
    c = vadd(c, vsc(V(0.0f, 0.0f, WALL_I), Np.z));
    c = vadd(c, vsc(V(WALL_I, WALL_I, 0.0f), Nn.x));
    c = vadd(c, vsc(V(0.0f, WALL_I, WALL_I), Nn.y));
    c = vadd(c, vsc(V(WALL_I, 0.0f, WALL_I), Nn.z));
    return c;
}

/* reflect_ray1 (:302-335) */
static V3 reflect_ray1(const Shade *c, V3 src, V3 dir) {
    Hdda hit = hdda(c, src, dir);
    V3 step = sign11v(dir);
    if (hit.state == OR_HIT) {
        V3 N = vnorm(vmul(vneg(step), hit.mask));
        float I = c->s->sun_color[3] * K_D * vdot(sun_neg(c), N);
        I = smax(0.0f, I);
        if (I != 0.0f && hdda(c, offset_origin(hit, step), sun_neg(c)).state == OR_HIT)
            return vadd(BASE_COLOR, vsc(sv(I, sun_rgb(c)), 0.05f));
        return vadd(BASE_COLOR, sv(I, sun_rgb(c)));
    }
    if (hit.state == OR_OOB) return walls(vnorm(vmul(vneg(step), hit.mask)));
    return dir;
}

/* reflect_ray2 (:244-300) */
static V3 reflect_ray2(const Shade *c, V3 src, V3 dir) {
    Hdda hit = hdda(c, src, dir);
    V3 step = sign11v(dir);
    if (hit.state == OR_HIT) {
        V3 N = vnorm(vmul(vneg(step), hit.mask));
        V3 rdir = vnorm(vsub(dir, vsc(sv(2.0f, N), vdot(dir, N))));
        V3 rsrc = offset_origin(hit, step);
        V3 rcol = reflect_ray1(c, rsrc, rdir);
        V3 mcol;
        float I = c->s->sun_color[3] * K_D * vdot(sun_neg(c), N);
        I = smax(0.0f, I);
        if (I != 0.0f && hdda(c, offset_origin(hit, step), sun_neg(c)).state == OR_HIT)
            mcol = vadd(BASE_COLOR, vsc(sv(I, sun_rgb(c)), 0.05f));
        else
            mcol = vadd(BASE_COLOR, sv(I, sun_rgb(c)));
        return vmix(mcol, rcol, REFLECTIVITY);
    }
    if (hit.state == OR_OOB) return walls(vnorm(vmul(vneg(step), hit.mask)));
    return dir;
}

static int on_boundary(V3 p, float s) {
    return fmodf(floorf(p.x), s) == 0.0f || fmodf(floorf(p.y), s) == 0.0f || fmodf(floorf(p.z), s) == 0.0f;
}

static float mask_dot(V3 m, V3 w) { return vdot(vmul(m, w), V(1.0f, 1.0f, 1.0f)); }

/* ray_trace (:152-242) */
static V3 ray_trace(const Shade *c, V3 src, V3 dir) {
    Hdda hit = hdda(c, src, dir);
    V3 step = sign11v(dir);
    const or_shade *s = c->s;
    if (hit.state == OR_HIT) {
        V3 grid = V(0.0f, 0.0f, 0.0f);
        if (s->show_345[2] == 1u && on_boundary(hit.p, 4096.0f)) grid = V(-0.3f, -0.3f, 1.0f);
        else if (s->show_345[1] == 1u && on_boundary(hit.p, 128.0f)) grid = V(0.6f, -0.2f, -0.2f);
        else if (s->show_345[0] == 1u && on_boundary(hit.p, 8.0f)) grid = V(-0.1f, 0.5f, 0.3f);
        switch (s->mode) {
        case 1:
            return vadd(vadds(grid, 0.1f), vmul(hit.mask, V(0.4f, 0.4f, 0.4f)));
        case 2: {
            float t = hit.i / 200.0f;
            return vadd(grid, vmix(V(0.72f, 1.0f, 0.99f), V(1.0f, 0.0f, 0.0f), t));
        }
        case 3: {
            V3 N = vnorm(vmul(vneg(step), hit.mask));
            float LN = smax(0.0f, s->sun_color[3] * vdot(sun_neg(c), N));
            V3 I_d = vsc(vmul(sv(K_D, sun_rgb(c)), BASE_COLOR), LN);
            V3 I_a = vmul(sv(K_A, AMBIENT_COLOR), BASE_COLOR);
            if (LN != 0.0f && hdda(c, offset_origin(hit, step), sun_neg(c)).state == OR_HIT) return I_a;
            return vadd(I_a, I_d);
        }
        case 4: {
            V3 N = vnorm(vmul(vneg(step), hit.mask));
            V3 mcol;
            float I = s->sun_color[3] * K_D * vdot(sun_neg(c), N);
            I = smax(0.0f, I);
            if (I != 0.0f && hdda(c, offset_origin(hit, step), sun_neg(c)).state == OR_HIT)
                mcol = vadd(BASE_COLOR, vsc(sv(I, sun_rgb(c)), 0.05f));
            else
                mcol = vadd(BASE_COLOR, sv(I, sun_rgb(c)));
            V3 rdir = vnorm(vsub(dir, vsc(sv(2.0f, N), vdot(dir, N))));
            V3 rsrc = offset_origin(hit, step);
            V3 rcol = reflect_ray2(c, rsrc, rdir);
            return vmix(mcol, rcol, REFLECTIVITY);
        }
        default:
            return vadds(grid, mask_dot(hit.mask, V(0.2f, 0.2f, 0.3f)));
        }
    }
    if (hit.state == OR_OOB) {
        switch (s->mode) {
        case 2: {
            float t = hit.i / 200.0f;
            return vadds(vmix(V(0.72f, 1.0f, 0.99f), V(1.0f, 0.0f, 0.0f), t), mask_dot(hit.mask, V(0.04f, 0.08f, 0.12f)));
        }
        case 4: {
            V3 N = vnorm(vmul(vneg(step), hit.mask));
            V3 Np = V(smax(0.0f, N.x), smax(0.0f, N.y), smax(0.0f, N.z));
            V3 Nn = vneg(V(smin(0.0f, N.x), smin(0.0f, N.y), smin(0.0f, N.z)));
            float t = hit.p.y / 4096.0f, w1 = WALL_I * 0.1f;
            V3 r = vsc(vmix(V(WALL_I, 0.0f, 0.0f), V(w1, 0.0f, 0.0f), t), Np.x);
            r = vadd(r, vsc(V(0.0f, WALL_I, 0.0f), Np.y));
            r = vadd(r, vsc(vmix(V(0.0f, 0.0f, WALL_I), V(0.0f, 0.0f, w1), t), Np.z));
            r = vadd(r, vsc(vmix(V(WALL_I, WALL_I, 0.0f), V(w1, w1, 0.0f), t), Nn.x));
            r = vadd(r, vsc(V(0.0f, WALL_I, WALL_I), Nn.y));
            r = vadd(r, vsc(vmix(V(WALL_I, 0.0f, WALL_I), V(w1, 0.0f, w1), t), Nn.z));
            return r;
        }
        default:
            return vadds(V(0.0f, 0.0f, 0.0f), mask_dot(hit.mask, V(0.01f, 0.02f, 0.03f)));
        }
    }
    return dir;
}

static uint32_t unorm8(float v) {
    float c = v > 0.0f ? (v < 1.0f ? v : 1.0f) : 0.0f;
    return (uint32_t)rintf(c * 255.0f);
}

void or_render_pixel(const or_layout *L, const float cam[16], uint32_t px, uint32_t py, float offset, float bound,
                     int clip, const or_shade *s, float rgb[3]) {
    Shade c = {L, bound, clip, s};
    float d[3];
    or_camera_ray(cam, px, py, offset, d);
    V3 col = ray_trace(&c, V(cam[0], cam[1], cam[2]), V(d[0], d[1], d[2]));
    rgb[0] = col.x, rgb[1] = col.y, rgb[2] = col.z;
}

typedef struct {
    const or_layout *L;
    const float *cam;
    uint32_t w, h;
    float offset, bound;
    int clip;
    const or_shade *s;
    float *rgba;
    uint32_t *rgba8;
    size_t n, next;
    pthread_mutex_t mu;
} RJob;

static void *render_worker(void *arg) {
    RJob *j = (RJob *)arg;
    for (;;) {
        pthread_mutex_lock(&j->mu);
        size_t b = j->next, e = b + 256 < j->n ? b + 256 : j->n;
        j->next = e;
        pthread_mutex_unlock(&j->mu);
        if (b >= e) break;
        for (size_t k = b; k < e; ++k) {
            float c[3];
            or_render_pixel(j->L, j->cam, (uint32_t)(k % j->w), (uint32_t)(k / j->w), j->offset, j->bound, j->clip,
                            j->s, c);
            if (j->rgba) {
                j->rgba[4 * k] = c[0], j->rgba[4 * k + 1] = c[1], j->rgba[4 * k + 2] = c[2];
                j->rgba[4 * k + 3] = 1.0f;
            }
            if (j->rgba8) j->rgba8[k] = unorm8(c[0]) | (unorm8(c[1]) << 8) | (unorm8(c[2]) << 16) | (255u << 24);
        }
    }
    return NULL;
}

void or_render_camera(const or_layout *L, const float cam[16], uint32_t w, uint32_t h, float offset, float bound,
                      int clip, const or_shade *s, float *rgba, uint32_t *rgba8, int nthreads) {
    RJob j;
    memset(&j, 0, sizeof j);
    j.L = L; j.cam = cam; j.w = w; j.h = h; j.offset = offset; j.bound = bound; j.clip = clip; j.s = s;
    j.rgba = rgba; j.rgba8 = rgba8; j.n = (size_t)w * h;
    if (nthreads < 1) nthreads = 1;
    pthread_mutex_init(&j.mu, NULL);
    pthread_t *t = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
    for (int i = 0; i < nthreads; ++i) pthread_create(&t[i], NULL, render_worker, &j);
    for (int i = 0; i < nthreads; ++i) pthread_join(t[i], NULL);
    free(t);
    pthread_mutex_destroy(&j.mu);
}

void or_trace_camera(const or_layout *L, const float cam[16], uint32_t w, uint32_t h, float offset, float bound,
                     int clip, or_hit *out, or_diag *diag, int nthreads) {
    Job j;
    memset(&j, 0, sizeof j);
    j.L = L; j.cam = cam; j.w = w; j.h = h; j.n = (size_t)w * h;
    j.offset = offset; j.bound = bound; j.clip = clip; j.out = out; j.diag = diag;
    run_job(&j, cam_worker, nthreads);
}

void or_trace_rays(const or_layout *L, const float *rays, size_t n, float bound, int clip, or_hit *out,
                   or_diag *diag, int nthreads) {
    Job j;
    memset(&j, 0, sizeof j);
    j.L = L; j.rays = rays; j.n = n; j.bound = bound; j.clip = clip; j.out = out; j.diag = diag;
    run_job(&j, ray_worker, nthreads);
}
