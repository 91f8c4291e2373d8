"""Allocation-free gridding entry point (GPU), through the C ABI.

tn_grid_static must give the same bits as tn_grid for the same plan and
inputs (the same kernels and summation order; only the workspace of the
partial sums differs): checked on two calls in a row, the second with a
smaller input (fewer samples and coils) whose workspace contents differ, on
a trajectory whose hot-spot cells have matrix rows longer than one part, so
the partial records go through the workspace, and for the 3D register-tile
gridding on split hot tiles. Without prepared launch tables it returns
TN_E_ARG and launches nothing.
"""

import ctypes

import numpy as np
import pytest
import torch

import tilenufft
from tilenufft import _core, _lib, testing

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _hot_coord(n, nbg, nhot, seed):
    rng = np.random.default_rng(seed)
    bg = rng.uniform(-n / 2, n / 2, (1, nbg, 2))
    hot = np.tile(np.array([[[0.3, -0.2]]]), (1, nhot, 1)) + rng.normal(0, 0.05, (1, nhot, 2))
    return np.concatenate([bg, hot], 1).astype(np.float32)


def _call(name, plan, k, nbatch):
    lib = _lib.load()
    out = torch.full((nbatch, *plan.os_shape), float("nan"), dtype=torch.complex64, device=DEV)
    rc = getattr(lib, name)(plan.interpolator.handle, _core.ptr(k), _core.ptr(out), nbatch, None,
                            _core.stream_ptr(DEV))
    _lib.check(rc, name)
    torch.cuda.synchronize()
    return out


def test_grid_static_matches_grid_on_two_calls():
    lib = _lib.load()
    cases = [(32, 1500, 3000, 16, 1), (24, 300, 900, 5, 2)]  # the second call is smaller
    for n, nbg, nhot, ncoil, seed in cases:
        print(f"tn_grid_static N={n} samples={nbg + nhot} coils={ncoil}", flush=True)
        coord = _hot_coord(n, nbg, nhot, seed)
        plan = tilenufft.fft.plan_nufft(coord, n, 4, 2.0, device=DEV, grid_mode=3, grid_split=256)
        k = testing.complex_normal((ncoil, plan.npts), seed=seed).to(DEV)
        with torch.cuda.device(DEV):
            rc = lib.tn_grid_static(plan.interpolator.handle, _core.ptr(k), _core.ptr(k), ncoil, None,
                                    _core.stream_ptr(DEV))
            assert rc == -1 and "tn_plan_prepare" in lib.tn_last_error().decode()
            plan.interpolator.prepare(ncoil)
            ref = _call("tn_grid", plan, k, ncoil)
            got = _call("tn_grid_static", plan, k, ncoil)
        assert not torch.isnan(ref.real).any()
        assert torch.equal(got.view(torch.float32), ref.view(torch.float32))


def test_grid_static_matches_grid_register_tiles():
    """The same for the 3D register-tile gridding: hot tiles split into
    parts of 64 entries, whose partial tiles go through the workspace; two
    calls in a row, the second smaller."""
    cases = [(16, 2000, 8, 3), (16, 700, 3, 4)]  # the second call is smaller
    for n, npts, ncoil, seed in cases:
        print(f"tn_grid_static tiles N={n} samples={npts} coils={ncoil}", flush=True)
        rng = np.random.default_rng(seed)
        coord = np.concatenate([rng.uniform(-n / 2, n / 2, (1, npts // 2, 3)),
                                rng.normal(0.0, 0.7, (1, npts - npts // 2, 3))], 1).astype(np.float32)
        plan = tilenufft.fft.plan_nufft(coord, n, 4, 2.0, device=DEV, grid_mode=4, grid_split=64)
        assert plan.interpolator.grid_mode_effective() == 4
        k = testing.complex_normal((ncoil, plan.npts), seed=seed).to(DEV)
        with torch.cuda.device(DEV):
            plan.interpolator.prepare(ncoil)
            ref = _call("tn_grid", plan, k, ncoil)
            got = _call("tn_grid_static", plan, k, ncoil)
        assert not torch.isnan(ref.real).any()
        assert torch.equal(got.view(torch.float32), ref.view(torch.float32))
