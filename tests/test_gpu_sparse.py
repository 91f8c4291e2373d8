"""GPU Cartesian sparse FFT vs the oracle and the reference's known answers
(tests/fft/test_sparse_fft.py). Bar: relative L2 <= 1e-5 (complex64) against
the oracle; the reference's known answers to its own 1e-4."""

import itertools

import numpy as np
import pytest
import torch

import tilenufft
from oracle import sparse_oracle as so
from tilenufft import testing

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
RTOL = 1e-5


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-30))


def full_indexes(ndim, ncontrasts, npix):
    nodes = np.arange(npix)
    mesh = np.meshgrid(*([nodes] * ndim), indexing="ij")
    idx = np.stack([m.ravel() for m in mesh], -1).astype(np.float32)[None]
    if ncontrasts > 1:
        idx = np.repeat(idx[None], ncontrasts, 0)
    return torch.from_numpy(idx)


@pytest.mark.parametrize("ndim, ncontrasts, lowrank", list(itertools.product([1, 2, 3], [1, 2], [False, True])))
def test_known_answers(ndim, ncontrasts, lowrank):
    if lowrank and ncontrasts == 1:
        pytest.skip("the reference's low-rank cases use several contrasts")
    npix, ncoils = 4, 2
    print(f"sparse known answers ndim={ndim} ncontrasts={ncontrasts} lowrank={lowrank}", flush=True)
    idx = full_indexes(ndim, ncontrasts, npix)
    lead = (ncoils,) if ncontrasts == 1 else (ncoils, ncontrasts)
    img = torch.zeros(lead + (npix,) * ndim, dtype=torch.complex64)
    img[(Ellipsis,) + (npix // 2,) * ndim] = 1.0
    eye = torch.eye(ncontrasts, dtype=torch.complex64) if lowrank else None
    k = tilenufft.fft.sparse_fft(img.clone(), idx, basis_adjoint=eye, device=DEV)
    ones = torch.ones(lead + (1, npix ** ndim), dtype=torch.complex64)
    np.testing.assert_allclose(k.cpu().numpy(), ones.numpy(), rtol=1e-4, atol=1e-4)
    back = tilenufft.fft.sparse_ifft(ones.clone(), idx, npix, basis=eye, device=DEV)
    np.testing.assert_allclose(back.cpu().numpy(), img.numpy(), rtol=1e-4, atol=1e-4)
    if ndim >= 2:
        kern = tilenufft.fft.plan_toeplitz_fft(idx, npix, basis=eye, device=DEV)
        out = tilenufft.fft.apply_sparse_fft_selfadj(img.clone(), kern, device=DEV)
        np.testing.assert_allclose(out.cpu().numpy(), img.numpy(), rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("shape, nframes", [((48, 40), 1), ((36, 45), 3), ((12, 10, 14), 2), ((64,), 2)])
def test_matches_oracle(shape, nframes):
    print(f"sparse vs oracle shape={shape} frames={nframes}", flush=True)
    rng = np.random.default_rng(len(shape) * 10 + nframes)
    ndim = len(shape)
    m = 300
    idx = np.stack([rng.integers(0, n, (nframes, 2, m)) for n in shape[::-1]], -1)
    idx[..., 1, :5, :] = idx[..., 0, :5, :]  # duplicated samples
    if nframes == 1:
        idx = idx[0]
    lead = (3,) if nframes == 1 else (3, nframes)
    x = testing.complex_normal(lead + shape, seed=1).numpy()
    w = rng.uniform(0.1, 1.0, idx.shape[:-1]).astype(np.float32)
    mask = tilenufft.fft.prepare_sampling(torch.from_numpy(idx), shape, device=DEV)
    omask = so.plan_sampling(idx, shape)
    y = tilenufft.fft.apply_sparse_fft(torch.from_numpy(x).to(DEV), mask, weight=torch.from_numpy(w)).cpu().numpy()
    y_ref = so.apply_sparse_fft(x, omask, weight=w)
    assert y.shape == y_ref.shape
    assert rel(y, y_ref) < RTOL
    k = testing.complex_normal(y.shape, seed=2).numpy()
    xa = tilenufft.fft.apply_sparse_ifft(torch.from_numpy(k).to(DEV), mask, weight=torch.from_numpy(w)).cpu().numpy()
    xa_ref = so.apply_sparse_ifft(k, omask, weight=w)
    assert xa.shape == xa_ref.shape
    assert rel(xa, xa_ref) < RTOL


@pytest.mark.parametrize("ndim", [2, 3])
def test_lowrank_and_toeplitz_match_oracle(ndim):
    print(f"sparse low-rank + toeplitz ndim={ndim}", flush=True)
    rng = np.random.default_rng(40 + ndim)
    shape = (32, 24) if ndim == 2 else (10, 12, 8)
    nf, nc = 6, 3
    idx = np.stack([rng.integers(0, n, (nf, 1, 200)) for n in shape[::-1]], -1)
    ba = (rng.standard_normal((nf, nc)) + 1j * rng.standard_normal((nf, nc))).astype(np.complex64)
    basis = np.ascontiguousarray(ba.conj().T)
    mask = tilenufft.fft.prepare_sampling(torch.from_numpy(idx), shape, device=DEV)
    omask = so.plan_sampling(idx, shape)
    x = testing.complex_normal((2, nc) + shape, seed=3).numpy()
    y = tilenufft.fft.apply_sparse_fft(torch.from_numpy(x).to(DEV), mask, basis_adjoint=ba).cpu().numpy()
    assert rel(y, so.apply_sparse_fft(x, omask, basis_adjoint=ba)) < RTOL
    k = testing.complex_normal(y.shape, seed=4).numpy()
    xa = tilenufft.fft.apply_sparse_ifft(torch.from_numpy(k).to(DEV), mask, basis=basis).cpu().numpy()
    assert rel(xa, so.apply_sparse_ifft(k, omask, basis=basis)) < RTOL
    for product in ("deepmr", "matvec"):
        kern = tilenufft.fft.plan_toeplitz_fft(torch.from_numpy(idx), shape, basis=torch.from_numpy(basis),
                                               device=DEV, lowrank_product=product)
        k_ref, lr = so.plan_toeplitz_fft(idx, shape, basis=basis)
        assert rel(kern.value.cpu().numpy(), k_ref) < RTOL
        out = tilenufft.fft.apply_sparse_fft_selfadj(torch.from_numpy(x).to(DEV), kern).cpu().numpy()
        assert rel(out, so.apply_sparse_fft_selfadj(x, k_ref, ndim, lr, product == "matvec")) < RTOL


def test_autograd_exact_vjp():
    print("sparse autograd", flush=True)
    rng = np.random.default_rng(7)
    shape = (20, 16)
    idx = np.stack([rng.integers(0, n, (1, 150)) for n in shape[::-1]], -1)
    mask = tilenufft.fft.prepare_sampling(torch.from_numpy(idx), shape, device=DEV)
    x = testing.complex_normal((2,) + shape, seed=8).to(DEV).requires_grad_(True)
    y = tilenufft.fft.apply_sparse_fft(x, mask)
    g = testing.complex_normal(tuple(y.shape), seed=9).to(DEV)
    (y.conj() * g).real.sum().backward()
    ref = float(np.prod(shape)) * tilenufft.fft.apply_sparse_ifft(g, mask)
    assert rel(x.grad.cpu().numpy(), ref.detach().cpu().numpy()) < 1e-5
