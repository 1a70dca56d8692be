"""Seeded random VDB345 trees for the parity tests: sphere unions over one or
several root nodes, value-noise isosurfaces, explicit node-4 active tiles and empty
leaves. The product tree is built from the primitives; the oracle tree gets the same
topology (every leaf with its mask, then the explicit nodes) and computes its own SDF
and layout."""
import numpy as np


def random_tree(wx, oracle_mod, rng):
    p = wx.Vdb345()
    centre = rng.uniform(-3000, 3000, size=3)
    n = int(rng.integers(1, 7))
    sp = np.column_stack([centre + rng.uniform(-400, 400, size=(n, 3)), rng.uniform(2, 60, size=n)])
    p.add_spheres(sp, threads=2)
    if rng.uniform() < 0.6:
        lo = (centre + rng.uniform(-300, 100, size=3)).astype(np.int32)
        hi = lo + rng.integers(40, 200, size=3).astype(np.int32)
        p.add_noise(int(rng.integers(1 << 30)), lo, hi, float(rng.uniform(0.01, 0.05)), float(rng.uniform(0.05, 0.3)),
                    threads=2)
    nodes = []
    if rng.uniform() < 0.5:  # an explicit node-4 with active tiles next to the geometry
        k4 = (np.floor(centre / 128).astype(np.int64) + rng.integers(-3, 4, size=3)).astype(np.int32)
        m = np.zeros(64, np.uint64)
        for b in rng.choice(4096, size=int(rng.integers(1, 40)), replace=False):
            m[b >> 6] |= np.uint64(1) << np.uint64(b & 63)
        nodes.append((4, k4, m))
    if rng.uniform() < 0.3:  # an empty leaf
        k3 = (np.floor(centre / 8).astype(np.int64) + rng.integers(-40, 40, size=3)).astype(np.int32)
        nodes.append((3, k3, np.zeros(8, np.uint64)))
    for lvl, k, m in nodes:
        p.add_nodes(lvl, k[None, :], m[None, :])
    p.compute_sdf(threads=2)
    # the oracle gets the same topology: every leaf (with its mask), then the explicit nodes
    o = oracle_mod.Vdb()
    org, masks = p.leaf_masks()
    for i in range(org.shape[0]):
        o.add_node(3, org[i], masks[i])
    for lvl, k, m in nodes:
        shift = {5: 12, 4: 7, 3: 3}[lvl]
        if lvl != 3:
            o.add_node(lvl, (k.astype(np.int64) << shift).astype(np.int32), m)
    o.compute_sdf()
    return p, o, centre
