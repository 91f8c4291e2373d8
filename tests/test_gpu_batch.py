"""Batched ragged launches and HIP-graph capture (GPU).

Each trajectory of a batch must give the same result as its own
single-plan NUFFT (the kernels are per sample / per tile, so interpolation
and gridding are bitwise equal; the batched FFT is held to 1e-6), and graph
replays must equal the eager result on every replay.
"""

import numpy as np
import pytest
import torch

import tilenufft
from oracle import nufft_oracle as orc
from tilenufft import testing
from tilenufft.batch import BatchedNUFFT

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def rel(a, b):
    a, b = torch.as_tensor(a), torch.as_tensor(b)
    return float(torch.linalg.norm(a - b) / torch.linalg.norm(b))


def _batch(ncoils=(3, 2, 4), n=48):
    counts = [5, 9, 7]
    coords = [testing.vd_spiral(L, 256, n, rotation=0.7 * i) for i, L in enumerate(counts)]
    plans = [tilenufft.fft.plan_nufft(c, n, 4, 2.0, device=DEV) for c in coords]
    b = BatchedNUFFT(plans, list(ncoils))
    g = torch.Generator().manual_seed(3)
    for k, c in enumerate(ncoils):
        b.image_view(k).copy_(testing.complex_normal((c, 1, n, n), generator=g).to(DEV))
    return b, coords, plans


def test_batch_matches_single_plan_and_oracle():
    print("batched forward/adjoint vs single-plan", flush=True)
    b, coords, plans = _batch()
    b.forward()
    b.adjoint()
    torch.cuda.synchronize()
    for k, p in enumerate(plans):
        img = b.image_view(k)[:, 0]
        y_single = tilenufft.fft.apply_nufft(img, p)
        y_batch = b.kspace_view(k)[:, 0].reshape(y_single.shape)
        assert rel(y_batch, y_single) < 1e-6
        y_ref = orc.nufft(img.cpu().numpy(), coords[k], 48, 4, 2.0)
        assert rel(y_batch.cpu(), torch.from_numpy(y_ref)) < 1e-5
        x_single = tilenufft.fft.apply_nufft_adj(y_batch, p)
        assert rel(b.image_view(k, adjoint=True)[:, 0], x_single) < 1e-6


def test_batch_kernels_bitwise_equal_to_single_plan():
    b, coords, plans = _batch()
    grid = torch.fft.fftn(torch.randn(b.grid.shape, dtype=torch.complex64, device=DEV), dim=(-2, -1))
    b.interp(grid)
    torch.cuda.synchronize()
    gvol = 96 * 96
    for k, p in enumerate(plans):
        c = b.ncoils[k]
        g_k = grid.reshape(-1)[b.grid_off[k]:b.grid_off[k] + c * gvol].view(c, 1, 96, 96)
        out = torch.empty((c, 1, p.npts), dtype=torch.complex64, device=DEV)
        p.interpolator.interp(g_k.contiguous(), out, c)
        assert torch.equal(out.view(torch.float32), b.kspace_view(k).contiguous().view(torch.float32))


def test_graph_replays_equal_eager():
    print("graph capture: forward+adjoint replays vs eager", flush=True)
    b, _, _ = _batch()
    b.forward()
    b.adjoint()
    torch.cuda.synchronize()
    k_ref = b.kspace.clone()
    x_ref = b.image_adj.clone()
    graph = b.capture()
    for r in range(3):
        b.kspace.zero_()
        b.image_adj.zero_()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        _ = (b.image * 2).sum().item()  # eager work between replays
        assert torch.equal(b.kspace, k_ref), f"replay {r} k-space differs"
        assert rel(b.image_adj, x_ref) < 1e-6, f"replay {r} image differs"


def test_autotune_picks_valid_config_and_keeps_results():
    b, coords, plans = _batch()
    b.forward()
    torch.cuda.synchronize()
    y0 = b.kspace.clone()
    res = tilenufft.autotune.tune_batch(plans, b.ncoils, reps=2, use_cache=False,
                                        candidates=dict(tiles=[(16, 16), (32, 32)], grid_chunks=[0, 2],
                                                        interp_chunks=[1, 4]))
    assert tuple(res["tile"]) in [(16, 16), (32, 32)]
    b2 = BatchedNUFFT(plans, b.ncoils)
    b2.image.copy_(b.image)
    b2.forward()
    torch.cuda.synchronize()
    assert torch.equal(b2.kspace, y0)  # interpolation does not depend on the tile
