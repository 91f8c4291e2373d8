"""bench.py's folding of rocprofv3 counter runs into bytes per stage launch
(CPU only): kernel-name shortening, stage assignment and the copy-kernel
calibration on synthetic counter files."""

import csv
import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


@pytest.mark.parametrize("name, short, stage", [
    ("void tn::(anonymous namespace)::grid_matrix_t_kernel<3, 2, 4>(tn::Job const*, tn::Item const*, long)",
     "grid_matrix_t_kernel<3, 2, 4>", "grid"),
    ("tn::(anonymous namespace)::samp_transpose_kernel<4>(tn::Job const*)", "samp_transpose_kernel<4>", "grid"),
    ("void tn::fft_rows_kernel<512, false>(float2 const*)", "fft_rows_kernel<512, false>", "fft_fwd"),
    ("void tn::fft_rows_kernel<512, true>(float2 const*)", "fft_rows_kernel<512, true>", "fft_adj"),
    ("interp_staged_kernel<2, 4>", "interp_staged_kernel<2, 4>", "interp"),
    ("void at::native::vectorized_elementwise_kernel<4, at::native::AUnaryFunctor<float, float, float, "
     "at::native::binary_internal::MulFunctor<float> >, std::array<char*, 2ul> >(int, "
     "at::native::AUnaryFunctor<float, float, float, at::native::binary_internal::MulFunctor<float> >, "
     "std::array<char*, 2ul>)", None, None),
])
def test_kernel_names_and_stages(name, short, stage):
    got = bench.short_kernel_name(name)
    if short is not None:
        assert got == short
    else:
        assert got.startswith("vectorized_elementwise_kernel<4")
    assert bench._stage_of(got) == stage


def _write_run(d, counter, rows):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "run_counter_collection.csv"), "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        for i, (k, v) in enumerate(rows):
            w.writerow([i, k, counter, v])


def test_fold_counters_scales_by_the_copy(tmp_path):
    mib = 64
    copy = "void at::native::vectorized_elementwise_kernel<4, Mul>(int)"
    grid = "void tn::grid_matrix_t_kernel<3, 2, 4>(int)"
    tr = "void tn::samp_transpose_kernel<4>(int)"
    # the copy reads as 0.5x its bytes in FETCH_SIZE and 1x in WRITE_SIZE (KiB)
    _write_run(tmp_path / "f", "FETCH_SIZE", [(copy, mib * 512), (copy, mib * 512), (copy, mib * 512),
                                              (grid, 1000), (grid, 1200), (tr, 300), ("void tn::other_kernel(int)", 9)])
    _write_run(tmp_path / "w", "WRITE_SIZE", [(copy, mib * 1024)] * 3 + [(grid, 2000), (tr, 100)])
    line = {"config": {"workload": "w"}, "kernels": {"variants": {"grid": "grid|mode=3", "interp": "interp|x"}}}
    (tmp_path / "b.log").write_text("bench: log line\n" + json.dumps(line) + "\n")
    out = bench.fold_counters([str(tmp_path / "f"), str(tmp_path / "w")], str(tmp_path / "b.log"), mib)
    assert out["fetch_scale"] == pytest.approx(2.0) and out["write_scale"] == pytest.approx(1.0)
    rec = out["stages"]["grid|mode=3"]
    assert rec["fetch_bytes"] == pytest.approx((1100 + 300) * 1024 * 2.0)
    assert rec["write_bytes"] == pytest.approx((2000 + 100) * 1024)
    assert out["variants"]["grid|mode=3"] == pytest.approx(rec["fetch_bytes"] + rec["write_bytes"])
    assert "interp|x" not in out["variants"]
