"""Host check of the shared-offset sliced layouts (csrc/slot_layout.cpp).

tests/slot_layout_check.cpp builds the full shared-offset layout and its
symmetric form (lower values read from their mirrors) for generated patterns
-- 3D and 2D boxes, a box with symmetric holes, a box with one mirror
missing, a periodic box, a small box, a random symmetric graph -- with the
library's own builders, and runs on the CPU what the FP64 and FP32 SpMV
kernels read: every index in range, and the CSR loop's bits
(Process/CG.f90:193-199) from every form; patterns the layouts cannot hold
are refused.
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
    out = os.path.join(REPO, "_build", "slot_layout_check")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    subprocess.run([cxx, "-O2", "-std=c++17", "-ffp-contract=off", "-I", CSRC, "-o", out,
                    os.path.join(REPO, "tests", "slot_layout_check.cpp"), os.path.join(CSRC, "slot_layout.cpp")],
                   check=True)
    return out


def test_slot_layouts_emulated_spmv(checker):
    r = subprocess.run([checker], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0 and "ALL OK" in r.stdout, r.stdout + r.stderr
    lines = {l.split()[1]: l for l in r.stdout.splitlines() if l.startswith("case ")}
    for name in ("box_20x17x13", "box_20x17x13_offdiag", "plane_40x37", "box_96x7x3", "box_with_holes"):
        assert "shared=1 sym=1" in lines[name], name
        f = dict(kv.split("=") for kv in lines[name].split()[2:6])
        assert int(f["stored"]) < 0.62 * int(f["full"]), name
    assert "shared=1 sym=0" in lines["box_missing_mirror"]
    for name in ("random_symmetric", "box_5x5x5", "box_periodic_x_offdiag"):
        assert "shared=0" in lines[name], name
