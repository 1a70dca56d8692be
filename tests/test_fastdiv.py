"""The traversal kernel divides p by the tile size with a three-operation
Markstein sequence instead of the IEEE division sequence (see div_tile in
woxel_amd/csrc/trace.hip). This pins that the two agree bit for bit on every
divisor the layout can produce, over the range of p the kernel uses it for."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("fastdiv") / "fastdiv_check")
    subprocess.run(["gcc", "-O2", "-mfma", "-ffp-contract=off", os.path.join(HERE, "fastdiv_check.c"), "-o", exe,
                    "-lm"], check=True)
    return exe


def run(exe, e0, e1, stride):
    r = subprocess.run([exe, str(e0), str(e1), str(stride)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


# biased exponents: 63 ~ 2^-64 (lower guard), 127 = [1, 2), 139 = [4096, 8192), 227 ~ 2^100 (upper guard)
@pytest.mark.parametrize("e0,e1,stride", [(63, 66, 4093), (127, 127, 509), (139, 140, 1021), (224, 227, 4093)])
def test_fast_tile_division_matches_ieee(checker, e0, e1, stride):
    out = run(checker, e0, e1, stride)
    assert "mismatches 0" in out


@pytest.mark.slow
def test_fast_tile_division_exhaustive_unit_binade(checker):
    # every significand of p in [1, 2) (both signs) for every divisor and scale;
    # correct rounding is scale-invariant while nothing under- or overflows
    assert "mismatches 0" in run(checker, 127, 127, 1)
