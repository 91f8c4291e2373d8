"""The C-ABI library and the host-side plan math (CPU only, no kernel launches).

- libtilenufft.so loads and exports every entry point include/tilenufft.h
  declares, and the ctypes mirror of tn_plan_params has the C layout.
- The product's host plan arithmetic agrees with the oracle's restatement.
"""

import ctypes
import os
import re
import subprocess

import numpy as np
import pytest
import torch

import tilenufft
from oracle import nufft_oracle as orc
from tilenufft import _core, _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tilenufft.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tn_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_symbols()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_lib.EXPORTED)
    assert "gfx950" in tilenufft.version()


@pytest.mark.parametrize("ctype", ["tn_plan_params", "tn_sense_job"])
def test_struct_layouts_match_c(tmp_path, ctype):
    """Every field of a ctypes mirror sits at the C offset (and sizeof matches)."""
    from tilenufft import sense

    P = {"tn_plan_params": _lib.PlanParams, "tn_sense_job": sense.SenseJob}[ctype]
    names = [f[0] for f in P._fields_]
    fmt = " ".join(["%zu"] * (len(names) + 1))
    args = ", ".join([f"sizeof({ctype})"] + [f"offsetof({ctype}, {n})" for n in names])
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "tilenufft.h"\n'
        f'int main(void){{printf("{fmt}\\n", {args});return 0;}}\n'
    )
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    assert [int(v) for v in out] == [ctypes.sizeof(P)] + [getattr(P, n).offset for n in names]


@pytest.mark.parametrize("w, osf", [(4, 2.0), (4, 1.5), (4, 1.25), (6, 2.0), (3, 1.5)])
def test_host_plan_math_matches_oracle(w, osf):
    beta = _core.kaiser_bessel_beta([w, w], [osf, osf])
    np.testing.assert_array_equal(beta, orc.kb_beta([w, w], [osf, osf]))
    s = _core.kernel_scale([w, w], beta)
    np.testing.assert_allclose(s, orc.kernel_scaling(beta, [w, w]), rtol=1e-12)
    for n in (15, 16, 128, 129):
        np.testing.assert_allclose(_core.apodization(n, np.float32(osf), w, beta[0]),
                                   orc.apodization_vector(n, np.float32(osf), w, beta[0]),
                                   rtol=3e-7)
    rng = np.random.default_rng(0)
    coord = rng.uniform(-40, 40, (3, 50, 2)).astype(np.float32)
    shape = np.asarray([80, 96], np.int16)
    ov = np.asarray([osf, osf], np.float32)
    np.testing.assert_array_equal(_core.scale_coordinates(coord, shape, ov),
                                  orc.scale_coord(coord, shape[::-1], ov[::-1]))
    np.testing.assert_array_equal(_core.oversampled_shape(shape, ov), orc.os_shape_of(shape, ov))


def test_cartesian_detection_matches_oracle():
    rng = np.random.default_rng(1)
    for coord in (rng.uniform(-8, 8, (20, 2)).astype(np.float32),
                  np.round(rng.uniform(-8, 8, (20, 3))).astype(np.float32),
                  np.stack([np.arange(10) * 0.5, np.arange(10)], -1).astype(np.float32)):
        ndim = coord.shape[-1]
        shape = np.asarray([16, 12, 8][:ndim], np.int16)
        assert (_core.cartesian_axes(coord, shape) == orc.cartesian_axes(coord, shape)).all()


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the no-GPU error path")
def test_no_gpu_fails_loudly():
    coord = np.zeros((4, 2), np.float32) + 0.3
    with pytest.raises(RuntimeError, match="HIP device"):
        tilenufft.fft.plan_nufft(coord, 8)
