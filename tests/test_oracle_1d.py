"""Pins for the oracle's 1D NUFFT (CPU only).

- the reference's 1D known answers (tests/fft/test_nufft.py:28-97): a centred
  delta maps to all ones and back on Cartesian coordinates (width 12 becomes
  the width-1 path), plain and low rank with an identity basis;
- the Kaiser-Bessel path against the FP64 direct NDFT, forward and adjoint;
- the dot-product identity <A x, y> = G <x, nufft_adj(y)>;
- 1D equals the 2D operator of a grid with a unit y axis (x-only
  coordinates), which reduces to it exactly in exact arithmetic: the 2D
  tables hold the same x taps and one y tap of weight 1;
- low rank equals the basis combination of the plain per-frame transforms.
"""

import itertools

import numpy as np
import pytest

from oracle import ndft
from oracle import nufft_oracle as orc


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def cn(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


@pytest.mark.parametrize("ncontrasts, ncoils, nslices", list(itertools.product([1, 2], [1, 2], [1, 2])))
def test_known_answers_1d(ncontrasts, ncoils, nslices):
    npix = 4
    coord = (np.arange(npix) - npix // 2).astype(np.float32)[None, :, None]  # (1, npix, 1)
    if ncontrasts > 1:
        coord = np.repeat(coord[None], ncontrasts, 0)
    lead = (nslices, ncoils) if ncontrasts == 1 else (nslices, ncoils, ncontrasts)
    img = np.zeros((*lead, npix), np.complex64)
    img[..., npix // 2] = 1
    k = orc.nufft(img, coord, None, width=12)
    np.testing.assert_allclose(k, np.ones((*lead, 1, npix)), rtol=1e-4, atol=1e-4)
    back = orc.nufft_adj(np.ones((*lead, 1, npix), np.complex64), coord, [npix], width=12)
    np.testing.assert_allclose(back, img, rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("ncontrasts", [2, 3])
def test_known_answers_lowrank_1d(ncontrasts):
    npix = 4
    coord = np.repeat((np.arange(npix) - npix // 2).astype(np.float32)[None, None, :, None], ncontrasts, 0)
    img = np.zeros((2, 2, ncontrasts, npix), np.complex64)
    img[..., npix // 2] = 1
    eye = np.eye(ncontrasts, dtype=np.complex64)
    k = orc.nufft(img, coord, None, width=12, basis_adjoint=eye)
    np.testing.assert_allclose(k, np.ones((2, 2, ncontrasts, 1, npix)), rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("osf, bound", [(2.0, 2e-3), (1.5, 8e-3), (1.25, 2e-2)])
def test_nufft1d_vs_ndft(osf, bound):
    rng = np.random.default_rng(11)
    n = 26
    coord = rng.uniform(-13, 13, (300, 1)).astype(np.float32)
    x = cn(rng, (n,))
    y = orc.nufft(x, coord, n, 4, osf)
    assert rel(y, ndft.ndft(x, coord, (n,)).reshape(y.shape)) < bound
    plan = orc.plan_nufft(coord, n, 4, osf)
    k = cn(rng, y.shape)
    xa = orc.apply_nufft_adj(k, plan)
    exact = ndft.ndft_adjoint(k.reshape(-1), coord, (n,)) / np.prod(plan["os_shape"])
    assert rel(xa, exact) < bound
    # a mirrored trajectory must not match
    assert rel(orc.nufft(x, -coord, n, 4, osf), ndft.ndft(x, coord, (n,)).reshape(y.shape)) > 0.1


def test_dot_product_identity_1d():
    rng = np.random.default_rng(12)
    coord = rng.uniform(-9, 9, (2, 3, 40, 1)).astype(np.float32)  # 2 frames
    plan = orc.plan_nufft(coord, 18, 5, 1.5)
    x = cn(rng, (3, 2, 18))
    y = orc.apply_nufft(x, plan)
    k = cn(rng, y.shape)
    lhs = np.vdot(y.astype(np.complex128), k)
    rhs = np.prod(plan["os_shape"]) * np.vdot(x.astype(np.complex128), orc.apply_nufft_adj(k, plan))
    assert abs(lhs - rhs) / abs(lhs) < 1e-5


def test_1d_equals_2d_with_a_unit_axis():
    rng = np.random.default_rng(13)
    n = 20
    c1 = rng.uniform(-10, 10, (2, 30, 1)).astype(np.float32)
    c2 = np.concatenate([c1, np.zeros_like(c1)], -1)  # (x, y = 0)
    p1 = orc.plan_nufft(c1, n, 4, 2.0)
    p2 = orc.plan_nufft(c2, (1, n), 4, 2.0)
    assert p2["os_shape"] == (1, p1["os_shape"][0]) and p2["width"] == (1, 4)
    x = cn(rng, (3, n))
    y1 = orc.apply_nufft(x, p1)
    y2 = orc.apply_nufft(x[:, None, :], p2)
    assert rel(y1, y2) < 1e-6
    k = cn(rng, y1.shape)
    assert rel(orc.apply_nufft_adj(k, p1), orc.apply_nufft_adj(k, p2)[:, 0]) < 1e-6


def test_lowrank_1d_equals_basis_combination():
    rng = np.random.default_rng(14)
    nf, nc, n = 4, 2, 16
    coord = rng.uniform(-8, 8, (nf, 1, 20, 1)).astype(np.float32)
    plan = orc.plan_nufft(coord, n, 4, 2.0)
    basis = cn(rng, (nc, nf))  # (ncoeff, nframes)
    ba = basis.conj().T.copy()  # (nframes, ncoeff)
    coef = cn(rng, (2, nc, n))
    y = orc.apply_nufft(coef, plan, basis_adjoint=ba)
    frames = np.einsum("fk,bkn->bfn", ba, coef).astype(np.complex64)
    assert rel(y, orc.apply_nufft(frames, plan)) < 1e-5
    k = cn(rng, y.shape)
    g = orc.apply_nufft_adj(k, plan, basis=basis)
    per = orc.apply_nufft_adj(k, plan)  # (2, nf, n)
    assert rel(g, np.einsum("kf,bfn->bkn", basis, per)) < 1e-5
