"""The host library (tree builder, .vdb reader, C ABI) under AddressSanitizer and
UBSan, built with g++ from its own sources (no GPU code): corrupted copies of the
reference's .vdb assets must be read or rejected cleanly, and random trees must go
through finalize and every query without a memory or undefined-behaviour error."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST_SOURCES = [os.path.join(ROOT, "woxel_amd", "csrc", f) for f in ("vdb_file.cpp", "vdb_host.cpp", "capi_host.cpp")]
FLAGS = ["-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
         "-fno-omit-frame-pointer", "-ffp-contract=off", "-pthread", "-I", os.path.join(ROOT, "include")]


@pytest.fixture(scope="module")
def exes(tmp_path_factory):
    """Both harnesses, built once (the host sources compiled once, in parallel)."""
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    d = tmp_path_factory.mktemp("sanitize")
    objs, procs = [], []
    for src in HOST_SOURCES + [os.path.join(ROOT, "tests", "sanitize", n + ".cpp") for n in ("fuzz_vdb", "fuzz_tree")]:
        o = str(d / (os.path.basename(src) + ".o"))
        objs.append(o)
        procs.append(subprocess.Popen(["g++", *FLAGS, "-c", src, "-o", o], stderr=subprocess.PIPE, text=True))
    for p in procs:
        err = p.communicate()[1]
        assert p.returncode == 0, err[-3000:]
    out = {}
    for k, name in enumerate(("fuzz_vdb", "fuzz_tree")):
        exe = str(d / name)
        r = subprocess.run(["g++", *FLAGS, "-o", exe, objs[len(HOST_SOURCES) + k], *objs[:len(HOST_SOURCES)]],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-3000:]
        out[name] = exe
    return out


def run(cmd):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0 and "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr, \
        (r.returncode, r.stderr[-4000:])
    return r.stdout


@pytest.mark.parametrize("asset,iters", [("icosahedron", 120), ("cube", 40)])
def test_corrupted_vdb_files(exes, tmp_path, asset, iters):
    out = run([exes["fuzz_vdb"], os.path.join(ROOT, "tests", "golden", asset + ".vdb"), str(iters),
               str(tmp_path / "c.vdb")])
    assert "rejected" in out


def test_random_trees(exes):
    assert "done" in run([exes["fuzz_tree"], "40"])
