"""The cell-record layout (voreen_amd/csrc/cell_layout.h): record index,
its inverse used by the record builder, and the neighbour steps of the
interior gradient, checked on the host for every cell of random volumes by
tests/native/cell_layout_check.cpp, compiled from the same header the
kernels include -- for the layout the library is built with (32-voxel
bricks, groups x-fastest) and for the alternatives the header offers (plain
x-fastest groups, other brick sizes, Morton order inside a brick).  CPU
only."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    pytest.skip("hipcc not found")


@pytest.mark.parametrize("brick_log2,morton", [(None, None), (4, 1), (0, 0), (3, 1), (2, 0)])
def test_cell_index_inverse_and_neighbours(tmp_path, brick_log2, morton):
    exe = str(tmp_path / "cell_layout_check")
    defs = [] if brick_log2 is None else [f"-DHIPDVR_CELL_BRICK_LOG2={brick_log2}", f"-DHIPDVR_CELL_MORTON={morton}"]
    subprocess.run([_hipcc(), "-x", "hip", "--cuda-host-only", "-O2", "-std=c++17", *defs,
                    "-I", os.path.join(ROOT, "voreen_amd", "csrc"),
                    os.path.join(ROOT, "tests", "native", "cell_layout_check.cpp"), "-o", exe],
                   check=True, capture_output=True)
    r = subprocess.run([exe, "60" if brick_log2 is None else "24"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    ok, checks = r.stdout.split()
    assert ok == "ok" and int(checks) > 1_000_000
