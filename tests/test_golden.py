"""Golden vectors (tests/golden/*.npz, written by tests/golden/make_golden.py
from the oracle): the oracle must reproduce them; the GPU path must match
them to 1e-5 relative L2."""

import glob
import os

import numpy as np
import pytest
import torch

from oracle import nufft_oracle as orc

FILES = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _load(path):
    d = np.load(path)  # allow_pickle is False by default
    shape = tuple(int(s) for s in d["shape"])
    shape = shape[0] if len(shape) == 1 else shape
    return d, shape


@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f) for f in FILES])
def test_oracle_reproduces_golden(path):
    d, shape = _load(path)
    y = orc.nufft(d["image"], d["coord"], shape, int(d["width"]), float(d["oversamp"]))
    xa = orc.nufft_adj(d["kspace"], d["coord"], shape, int(d["width"]), float(d["oversamp"]))
    assert rel(y, d["nufft"]) < 1e-6
    assert rel(xa, d["nufft_adj"]) < 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f) for f in FILES])
def test_gpu_matches_golden(path):
    import tilenufft

    d, shape = _load(path)
    print(f"golden {os.path.basename(path)}", flush=True)
    plan = tilenufft.fft.plan_nufft(d["coord"], shape, int(d["width"]), float(d["oversamp"]),
                                    device="cuda:0")
    y = tilenufft.fft.apply_nufft(torch.from_numpy(d["image"]).cuda(), plan).cpu().numpy()
    xa = tilenufft.fft.apply_nufft_adj(torch.from_numpy(d["kspace"]).cuda(), plan).cpu().numpy()
    assert rel(y, d["nufft"]) < 1e-5
    assert rel(xa, d["nufft_adj"]) < 1e-5
