"""Host-built gridding work lists (CPU only, no device).

The tile-matrix gridding items are built on the host from the plan's
per-tile batch counts (deep-mr_amd/csrc/tn_worklist.h). A host-only program
compiled from that header with g++ checks, over 400 random tile / batch /
coil / part-size configurations, that every (tile, coil record) is covered
by items whose batch ranges partition its bin in part order, that split
tiles have one fold item per slot with the right part count, and that the
fold slots' scratch regions hold every part's partial tile without overlap
(the kernel writes part q of slot s at slot_base[s] + q * 2 * coils * cells).
"""

import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_tile_matrix_worklist_invariants(tmp_path):
    exe = tmp_path / "worklist_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "deep-mr_amd", "csrc"),
                    os.path.join(ROOT, "tests", "native", "worklist_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip() == "ok 400", out.stdout
