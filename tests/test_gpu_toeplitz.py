"""Toeplitz normal operator on the GPU vs the oracle, and the reference's
self-adjoint known answers (tests/fft/test_nufft.py: delta -> delta)."""

import itertools

import numpy as np
import pytest
import torch

import tilenufft
from oracle import nufft_oracle as orc
from tilenufft import testing

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.mark.parametrize("ndim, ncontrasts", list(itertools.product([2, 3], [1, 2])))
def test_selfadjoint_delta_known_answer(ndim, ncontrasts):
    npix, ncoils = 4, 2
    print(f"toeplitz delta ndim={ndim} ncontrasts={ncontrasts}", flush=True)
    nodes = np.arange(npix) - npix // 2
    mesh = np.meshgrid(*([nodes] * ndim), indexing="ij")
    coord = np.stack([m.ravel() for m in mesh], -1).astype(np.float32)[None]
    if ncontrasts > 1:
        coord = np.repeat(coord[None], ncontrasts, 0)
    lead = (ncoils,) if ncontrasts == 1 else (ncoils, ncontrasts)
    img = torch.zeros((*lead, *(npix,) * ndim), dtype=torch.complex64)
    img[(Ellipsis,) + (npix // 2,) * ndim] = 1.0
    kern = tilenufft.fft.plan_toeplitz_nufft(torch.from_numpy(coord), npix, width=12, device=DEV)
    out = tilenufft.fft.apply_nufft_selfadj(img.clone(), kern, device=DEV)
    np.testing.assert_allclose(out.numpy(), img.numpy(), rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("ndim", [2, 3])
def test_toeplitz_matches_oracle(ndim):
    n = 32 if ndim == 2 else 12
    print(f"toeplitz vs oracle ndim={ndim}", flush=True)
    coord = testing.golden_angle_radial(20, 64, n) if ndim == 2 else testing.kooshball(40, 24, n)
    x = testing.complex_normal((2,) + (n,) * ndim, seed=3).numpy()
    kern = tilenufft.fft.plan_toeplitz_nufft(coord, n, width=4, device=DEV)
    ref_k, kshape = orc.plan_toeplitz(coord, n, width=4)
    assert tuple(kern.shape) == kshape
    assert rel(kern.value.cpu().numpy(), ref_k) < 1e-5
    y = tilenufft.fft.apply_nufft_selfadj(torch.from_numpy(x).to(DEV), kern).cpu().numpy()
    assert rel(y, orc.apply_selfadj(x, ref_k, kshape)) < 1e-5
    op = tilenufft.linops.NUFFTGramOp(torch.from_numpy(coord), n, width=4, device=DEV)
    assert rel(op(torch.from_numpy(x).to(DEV)).cpu().numpy(), y) < 1e-6


@pytest.mark.parametrize("kind, product", list(itertools.product(["complex", "real"], ["deepmr", "matvec"])))
def test_lowrank_toeplitz_matches_oracle(kind, product):
    n, nf, nc = 24, 6, 3
    print(f"low-rank toeplitz basis={kind} product={product}", flush=True)
    coord = np.stack([testing.vd_spiral(2, 128, n, rotation=f * testing.GOLDEN_ANGLE) for f in range(nf)])
    rng = np.random.default_rng(31)
    if kind == "complex":
        basis = (rng.standard_normal((nc, nf)) + 1j * rng.standard_normal((nc, nf))).astype(np.complex64)
    else:
        basis = rng.standard_normal((nc, nf)).astype(np.float32)
    kern = tilenufft.fft.plan_toeplitz_nufft(torch.from_numpy(coord), n, basis=torch.from_numpy(basis),
                                             device=DEV, lowrank_product=product)
    k_ref, ks = orc.plan_toeplitz(coord, n, basis=basis)
    assert kern.islowrank and tuple(kern.value.shape) == k_ref.shape
    assert rel(kern.value.cpu().numpy(), k_ref) < 1e-5
    x = testing.complex_normal((2, nc, n, n), seed=32)
    out = tilenufft.fft.apply_nufft_selfadj(x.to(DEV), kern).cpu().numpy()
    ref = orc.apply_selfadj(x.numpy(), k_ref, ks, lowrank=True, matvec=product == "matvec")
    assert rel(out, ref) < 1e-5
