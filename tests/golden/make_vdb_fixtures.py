"""Copy the reference's .vdb test assets (the files its reader tests open,
src/vdb/read.rs:717-744) into this folder as fixtures, with their SHA-256.

The files are data (OpenVDB level sets of a cube and an icosahedron); the
known answers that pin a reader on them are the reference test's assertion
(read.rs:770-802): active leaf voxels == the grid's file_voxel_count metadata.
Run from the repository root: python tests/golden/make_vdb_fixtures.py
"""
import hashlib
import json
import os
import shutil

SRC = "/root/reference/assets"
DST = os.path.dirname(os.path.abspath(__file__))

out = {}
for name in ("cube", "icosahedron"):
    src = os.path.join(SRC, f"{name}.vdb")
    dst = os.path.join(DST, f"{name}.vdb")
    shutil.copyfile(src, dst)
    out[f"{name}.vdb"] = {"sha256": hashlib.sha256(open(dst, "rb").read()).hexdigest(), "grid": f"ls_{name}"}
json.dump(out, open(os.path.join(DST, "vdb_fixtures.json"), "w"), indent=1)
print(json.dumps(out, indent=1))
