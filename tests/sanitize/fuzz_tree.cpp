// Host-code sanitizer harness (built by tests/test_host_sanitizers.py with
// AddressSanitizer and UBSan): random trees -- voxels, spheres, noise, explicit nodes
// with active tiles, coordinates up to +-2^23 -- through finalize and every query.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>
#include "woxel_hip.h"

int main(int argc, char **argv) {
    const int iters = argc > 1 ? std::atoi(argv[1]) : 50;
    std::mt19937_64 rng(777);
    for (int it = 0; it < iters; ++it) {
        wx_vdb *v = nullptr;
        wx_vdb_create(&v);
        const int span = 1 << (4 + it % 20);
        auto c = [&] { return (int32_t)((int64_t)(rng() % (2ull * span)) - span); };
        for (int k = 0; k < 200; ++k) wx_vdb_set_voxel(v, c(), c(), c(), 0x3F800000u);
        double sp[8] = {(double)c(), (double)c(), (double)c(), 5.0 + (double)(rng() % 20), (double)c(), (double)c(), (double)c(), 3.0};
        int32_t lo[3] = {-span, -span, -span}, hi[3] = {span, span, span};
        wx_vdb_add_spheres(v, sp, 2, lo, hi, 2);
        if (it % 3 == 0) {
            int32_t nlo[3] = {c(), c(), c()}, nhi[3] = {nlo[0] + 40, nlo[1] + 33, nlo[2] + 57};
            float thr;
            wx_vdb_add_noise(v, rng(), nlo, nhi, 0.05f, 0.2f, 2, &thr);
        }
        if (it % 4 == 1) {  // explicit nodes with active tiles
            std::vector<uint64_t> m5(512), m4(64), m3(8);
            for (auto &w : m5) w = rng() & rng() & rng();
            for (auto &w : m4) w = rng() & rng();
            int32_t k5[3] = {c() >> 12, c() >> 12, c() >> 12}, k4[3] = {c() >> 7, c() >> 7, c() >> 7}, k3[3] = {c() >> 3, c() >> 3, c() >> 3};
            wx_vdb_add_nodes(v, 5, k5, m5.data(), 1);
            wx_vdb_add_nodes(v, 4, k4, m4.data(), 1);
            wx_vdb_add_nodes(v, 3, k3, m3.data(), 1);
        }
        int rc = wx_vdb_finalize(v, 3);
        if (rc == WX_OK) {
            uint64_t n[3];
            wx_vdb_count_nodes(v, n);
            std::vector<uint32_t> k5(n[0] * 1024), v5(n[0] * 1024), k4(n[1] * 128), v4(n[1] * 128), v3(n[2] * 16);
            wx_vdb_ref_masks(v, k5.data(), v5.data(), k4.data(), v4.data(), v3.data());
            std::vector<uint32_t> a5(n[0] * 32768), a4(n[1] * 4096), a3(n[2] * 512);
            wx_vdb_ref_atlas(v, a5.data(), a4.data(), a3.data());
            std::vector<int32_t> org(n[2] * 3);
            wx_vdb_leaf_origins(v, org.data());
            for (int q = 0; q < 2000; ++q) {
                int kind, lvl;
                uint64_t val;
                wx_vdb_get_voxel(v, c(), c(), c(), &kind, &val, &lvl);
            }
            wx_grid_desc d;
            wx_vdb_layout(v, &d);
        }
        wx_vdb_destroy(v);
        if (it % 10 == 0) std::printf("iter %d span %d rc %d\n", it, span, rc);
    }
    std::printf("done\n");
    return 0;
}
