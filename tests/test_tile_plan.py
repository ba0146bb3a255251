"""Host check of the tiled DIC sweeps' plan (csrc/tile_plan.cpp).

tests/tile_plan_check.cpp builds the plan for generated patterns -- natural
boxes in 3D and 2D with and without an internal permutation, a random sorted
symmetric pattern whose levels overfill a tile key (the split path), and a
randomly renumbered box, which must be refused -- and runs on the CPU, in
the kernels' order, what the device sweeps do with it: every code in range,
every value read only after it is made, every wavefront schedule covering
each tile once in increasing order, and pivots and sweeps bitwise equal to
the sequential row-order ones (Prec1.f90:51-59, CG.f90:135-156).
"""
import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(REPO, "t-flows-old-compressible_amd", "csrc")


@pytest.fixture(scope="module")
def checker():
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    out = os.path.join(REPO, "_build", "tile_plan_check")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    subprocess.run([cxx, "-O2", "-std=c++17", "-pthread", "-I", CSRC, "-o", out,
                    os.path.join(REPO, "tests", "tile_plan_check.cpp"), os.path.join(CSRC, "tile_plan.cpp")],
                   check=True)
    return out


def test_tile_plan_emulated_sweeps(checker):
    r = subprocess.run([checker], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0 and "ALL OK" in r.stdout, r.stdout + r.stderr
    lines = {l.split()[1]: l for l in r.stdout.splitlines() if l.startswith("case ")}
    assert "plan=1 box=1" in lines["box_20x17x13"]
    assert "plan=1 box=1" in lines["box_20x17x13_internal_permutation"]
    assert "plan=1" in lines["random_sorted_symmetric"] and "splits=0" not in lines["random_sorted_symmetric"]
    assert "plan=0" in lines["box_randomly_renumbered"]
