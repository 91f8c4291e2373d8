"""Host-built gridding work lists (CPU only, no device).

The matrix-gridding items are built on the host from the plan's CSR row
offsets (deep-mr_amd/csrc/tn_worklist.h). A host-only program compiled from
that header with g++ checks, over 400 random row-length / threshold / cell
limit / coil-record configurations, that the runs cover every cell exactly
once (long rows first; all-empty matrices of sample-less plans included),
that short runs respect the kernel's cells per item
and LDS entry stage and carry the entry range of their rows, and that long
rows are split into in-order parts whose partial records are disjoint and
inside the scratch size, with one fold item each.
"""

import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_matrix_worklist_invariants(tmp_path):
    exe = tmp_path / "worklist_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "deep-mr_amd", "csrc"),
                    os.path.join(ROOT, "tests", "native", "worklist_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip() == "ok 400", out.stdout
