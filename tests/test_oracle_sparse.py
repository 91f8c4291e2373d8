"""Pins for the oracle's Cartesian sparse FFT (CPU only).

- the reference's known answers (tests/fft/test_sparse_fft.py): a centred
  delta transforms to all ones, all ones back to the delta, and the
  self-adjoint operator maps the delta to itself, with and without an
  identity subspace basis, in 1D/2D/3D with contrasts, coils and slices;
- the FP64 direct DFT at the integer frequencies idx - N//2 (odd and even
  sizes, duplicated samples), and its adjoint for the scatter;
- a general complex basis equals the basis-weighted per-frame operators.
"""

import itertools

import numpy as np
import pytest

from oracle import ndft
from oracle import sparse_oracle as so

TOL = 1e-4


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def cn(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def full_indexes(ndim, ncontrasts, npix):
    """All grid points, single shot, x first (the reference's generator)."""
    nodes = np.arange(npix)
    mesh = np.meshgrid(*([nodes] * ndim), indexing="ij")
    idx = np.stack([m.ravel() for m in mesh], -1).astype(np.float32)[None]
    if ncontrasts > 1:
        idx = np.repeat(idx[None], ncontrasts, 0)
    return idx


def lead_shape(ndim, ncontrasts, ncoils, nslices):
    lead = (nslices, ncoils) if ndim < 3 else (ncoils,)
    return lead if ncontrasts == 1 else lead + (ncontrasts,)


@pytest.mark.parametrize("ndim, ncontrasts, ncoils, nslices",
                         list(itertools.product([1, 2, 3], [1, 2], [1, 2], [1, 2])))
def test_known_answers(ndim, ncontrasts, ncoils, nslices):
    npix = 4
    if ndim == 3 and nslices > 1:
        pytest.skip("the reference has no slice axis in 3D")
    idx = full_indexes(ndim, ncontrasts, npix)
    lead = lead_shape(ndim, ncontrasts, ncoils, nslices)
    img = np.zeros(lead + (npix,) * ndim, np.complex64)
    img[(Ellipsis,) + (npix // 2,) * ndim] = 1.0
    k = so.sparse_fft(img, idx)
    np.testing.assert_allclose(k, np.ones(lead + (1, npix ** ndim)), rtol=TOL, atol=TOL)
    back = so.sparse_ifft(np.ones(lead + (1, npix ** ndim), np.complex64), idx, npix)
    np.testing.assert_allclose(back, img, rtol=TOL, atol=TOL)
    if ndim >= 2:
        kern, lr = so.plan_toeplitz_fft(idx, npix)
        out = so.apply_sparse_fft_selfadj(img, kern, ndim, lr)
        np.testing.assert_allclose(out, img, rtol=TOL, atol=TOL)


@pytest.mark.parametrize("ndim, ncontrasts, ncoils", list(itertools.product([1, 2, 3], [2, 3], [1, 2])))
def test_known_answers_lowrank_identity(ndim, ncontrasts, ncoils):
    npix = 4
    idx = full_indexes(ndim, ncontrasts, npix)
    eye = np.eye(ncontrasts, dtype=np.complex64)
    lead = (ncoils, ncontrasts)
    img = np.zeros(lead + (npix,) * ndim, np.complex64)
    img[(Ellipsis,) + (npix // 2,) * ndim] = 1.0
    k = so.sparse_fft(img, idx, basis_adjoint=eye)
    np.testing.assert_allclose(k, np.ones(lead + (1, npix ** ndim)), rtol=TOL, atol=TOL)
    back = so.sparse_ifft(np.ones(lead + (1, npix ** ndim), np.complex64), idx, npix, basis=eye)
    np.testing.assert_allclose(back, img, rtol=TOL, atol=TOL)
    if ndim >= 2:
        kern, lr = so.plan_toeplitz_fft(idx, npix, basis=eye)
        for matvec in (False, True):
            out = so.apply_sparse_fft_selfadj(img, kern, ndim, lr, matvec)
            np.testing.assert_allclose(out, img, rtol=TOL, atol=TOL)


@pytest.mark.parametrize("shape", [(6, 7), (5, 8, 6), (9,)])
def test_matches_direct_dft(shape):
    rng = np.random.default_rng(len(shape))
    ndim = len(shape)
    m = 40
    idx = np.stack([rng.integers(0, n, m) for n in shape[::-1]], -1)  # x first
    idx[5] = idx[3]  # a duplicated sample
    idx = idx[None, None]  # (1, 1, m, ndim)
    x = cn(rng, (2,) + shape)
    coord = (idx - np.asarray([n // 2 for n in shape[::-1]])).astype(np.float64)
    y = so.sparse_fft(x, idx)
    y_ref = ndft.ndft(x, coord.reshape(-1, ndim), shape).reshape(y.shape)
    assert rel(y, y_ref) < 1e-5
    k = cn(rng, y.shape)
    xa = so.sparse_ifft(k, idx, shape)
    xa_ref = ndft.ndft_adjoint(k.reshape(2, -1), coord.reshape(-1, ndim), shape) / np.prod(shape)
    assert rel(xa, xa_ref) < 1e-5


def test_lowrank_equals_basis_combination():
    rng = np.random.default_rng(5)
    nf, nc, shape = 5, 2, (8, 6)
    idx = np.stack([rng.integers(0, n, (nf, 1, 30)) for n in shape[::-1]], -1)
    ba = cn(rng, (nf, nc))
    x = cn(rng, (3, nc) + shape)
    y = so.sparse_fft(x, idx, basis_adjoint=ba)
    xf = np.einsum("fc,bc...->bf...", ba, x).astype(np.complex64)
    assert rel(y, so.sparse_fft(xf, idx)) < 1e-5
    basis = cn(rng, (nc, nf))
    k = cn(rng, (3, nf, 1, 30))
    xa = so.sparse_ifft(k, idx, shape, basis=basis)
    ref = np.einsum("cf,bf...->bc...", basis, so.sparse_ifft(k, idx, shape))
    assert rel(xa, ref) < 1e-5
