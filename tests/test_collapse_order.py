"""The wide-node collapse (CollapseWideBvh, csrc/bvh_build.cpp) on the host,
below and above the size from which it collapses subtrees on host threads:
nodes numbered in the order of the serial depth-first collapse, every item in
exactly one leaf, every child box the union of what lies below it.  The
checker (collapse_check.cpp) is built with the library's own compiler and
flags; no GPU is used."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc is not installed")
def test_collapse_numbering_items_and_boxes(tmp_path):
    exe = str(tmp_path / "collapse_check")
    csrc = os.path.join(ROOT, "azimuth_amd", "csrc")
    cmd = [HIPCC, "-std=c++17", "-O2", "-x", "hip", "--offload-arch=gfx950", "-ffp-contract=off",
           "-I" + csrc, "-I" + os.path.join(ROOT, "include"),
           os.path.join(csrc, "bvh_build.cpp"), os.path.join(HERE, "collapse_check.cpp"), "-o", exe]
    build = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert build.returncode == 0, build.stderr[-2000:]
    run = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    assert run.returncode == 0 and run.stdout.strip().endswith("ok"), run.stdout[-2000:] + run.stderr[-2000:]
