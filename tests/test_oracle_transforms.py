"""Pins for the oracle's full transforms (CPU only).

- resize: the reference's known answers (tests/signal/test_resize.py:9-91).
- Cartesian width-1 path and FFT conventions: the reference's known answers
  (tests/fft/test_nufft.py, centred delta <-> all ones).
- Kaiser-Bessel path: error against the FP64 direct NDFT (the closed form the
  NUFFT approximates), on anisotropic shapes so swapped axes fail, and the
  dot-product identity <A x, y> = prod(os_shape) <x, nufft_adj(y)>.
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


# --- resize known answers (tests/signal/test_resize.py) ---------------------
def test_resize_known_answers():
    x = np.array([[0, 0, 0], [0, 1, 0], [0, 0, 0]])
    np.testing.assert_array_equal(orc.resize(x, [5, 5]), np.pad(x, 1))
    np.testing.assert_array_equal(orc.resize(x, [4, 4]),
                                  [[0, 0, 0, 0], [0, 0, 0, 0], [0, 0, 1, 0], [0, 0, 0, 0]])
    np.testing.assert_array_equal(orc.resize(x[None], [5, 5])[0], np.pad(x, 1))
    y = np.pad(x, 1)
    np.testing.assert_array_equal(orc.resize(y, [3, 3]), x)
    z = np.array([[0, 0, 0, 0], [0, 0, 0, 0], [0, 0, 1, 0], [0, 0, 0, 0]])
    np.testing.assert_array_equal(orc.resize(z, [3, 3]), x)
    np.testing.assert_array_equal(orc.resize(z[None], [3, 3])[0], x)


# --- Cartesian known answers (tests/fft/test_nufft.py) ------------------------
def _cart_coord(ndim, ncontrasts, npix):
    nodes = np.arange(npix) - npix // 2
    mesh = np.meshgrid(*([nodes] * ndim), indexing="ij")
    coord = np.stack([m.ravel() for m in mesh], -1).astype(np.float32)[None]
    if ncontrasts > 1:
        coord = np.repeat(coord[None], ncontrasts, 0)
    return coord


@pytest.mark.parametrize("ndim, ncontrasts, ncoils, nslices",
                         list(itertools.product([2, 3], [1, 2], [1, 2], [1, 2])))
def test_cartesian_delta_ones(ndim, ncontrasts, ncoils, nslices):
    npix = 4
    coord = _cart_coord(ndim, ncontrasts, npix)
    lead = (nslices, ncoils) if ndim == 2 else (ncoils,)
    if ncontrasts > 1:
        lead = lead + (ncontrasts,)
    img = np.zeros((*lead, *(npix,) * ndim), np.complex64)
    img[(Ellipsis,) + (npix // 2,) * ndim] = 1
    k = orc.nufft(img, coord, None, width=12)
    np.testing.assert_allclose(k, np.ones((*lead, 1, npix**ndim)), rtol=1e-4, atol=1e-4)
    back = orc.nufft_adj(np.ones((*lead, 1, npix**ndim), np.complex64), coord, [npix] * ndim, width=12)
    np.testing.assert_allclose(back, img, rtol=1e-4, atol=1e-4)


# --- Kaiser-Bessel path vs the FP64 NDFT ---------------------------------------
@pytest.mark.parametrize("osf, bound", [(2.0, 2e-3), (1.5, 8e-3), (1.25, 2e-2)])
def test_nufft2d_vs_ndft_anisotropic(osf, bound):
    rng = np.random.default_rng(1)
    shape = (20, 28)  # (ny, nx)
    coord = np.stack([rng.uniform(-14, 14, 300), rng.uniform(-10, 10, 300)], -1).astype(np.float32)
    x = cn(rng, shape)
    y = orc.nufft(x, coord, shape, 4, osf)
    exact = ndft.ndft(x, coord, shape).reshape(y.shape)
    assert rel(y, exact) < bound
    k = cn(rng, y.shape)
    plan = orc.plan_nufft(coord, shape, 4, osf)
    xa = orc.apply_nufft_adj(k, plan)
    exact_a = ndft.ndft_adjoint(k.reshape(-1), coord, shape) / np.prod(plan["os_shape"])
    assert rel(xa, exact_a) < bound
    # swapping x/y in the coordinates must break the match
    swapped = orc.nufft(x, coord[:, ::-1].copy(), shape, 4, osf)
    assert rel(swapped, exact) > 0.1


@pytest.mark.parametrize("osf, bound", [(2.0, 3e-3), (1.5, 1e-2)])
def test_nufft3d_vs_ndft(osf, bound):
    rng = np.random.default_rng(2)
    shape = (8, 10, 12)
    coord = np.stack([rng.uniform(-6, 6, 200), rng.uniform(-5, 5, 200), rng.uniform(-4, 4, 200)],
                     -1).astype(np.float32)
    x = cn(rng, shape)
    y = orc.nufft(x, coord, shape, 4, osf)
    assert rel(y, ndft.ndft(x, coord, shape).reshape(y.shape)) < bound


def test_dot_product_identity():
    rng = np.random.default_rng(3)
    for ndim, shape in ((2, (18, 22)), (3, (8, 8, 10))):
        coord = rng.uniform(-4, 4, (2, 50, ndim)).astype(np.float32)  # 2 frames
        plan = orc.plan_nufft(coord, shape, 4, 1.5)
        x = cn(rng, (3, 2, *shape))
        y = orc.apply_nufft(x, plan)
        k = cn(rng, y.shape)
        xa = orc.apply_nufft_adj(k, plan)
        lhs = np.vdot(y.astype(np.complex128), k)
        rhs = np.prod(plan["os_shape"]) * np.vdot(x.astype(np.complex128), xa)
        assert abs(lhs - rhs) / abs(lhs) < 1e-5


def test_interp_grid_are_transposes():
    rng = np.random.default_rng(4)
    coord = rng.uniform(0, 30, (3, 1, 40, 2)).astype(np.float32)
    interp = orc.plan_interpolator(coord, [30, 30], [4, 4], [np.float32(6.0)] * 2)
    g = cn(rng, (3, 2, 30, 30))
    s = cn(rng, (3, 2, 40))
    lhs = np.vdot(orc.interpolate_raw(g, interp).astype(np.complex128), s)
    rhs = np.vdot(g.astype(np.complex128), orc.grid_raw(s, interp))
    assert abs(lhs - rhs) / abs(lhs) < 1e-6


def test_weight_applies_after_interp_and_before_grid():
    rng = np.random.default_rng(5)
    coord = rng.uniform(-8, 8, (30, 2)).astype(np.float32)
    plan = orc.plan_nufft(coord, 16, 4, 2.0)
    x = cn(rng, (16, 16))
    w = rng.uniform(0.2, 2.0, 30).astype(np.float32)
    np.testing.assert_allclose(orc.apply_nufft(x, plan, weight=w), w * orc.apply_nufft(x, plan),
                               rtol=1e-6)
    k = cn(rng, (1, 30))
    np.testing.assert_allclose(orc.apply_nufft_adj(k, plan, weight=w),
                               orc.apply_nufft_adj(w * k, plan), rtol=1e-6, atol=1e-7)


def test_lowrank_with_identity_basis_equals_plain():
    rng = np.random.default_rng(6)
    nf = 3
    coord = rng.uniform(-6, 6, (nf, 1, 25, 2)).astype(np.float32)
    plan = orc.plan_nufft(coord, 12, 4, 2.0)
    x = cn(rng, (2, nf, 12, 12))
    eye = np.eye(nf, dtype=np.complex64)
    np.testing.assert_allclose(orc.apply_nufft(x, plan, basis_adjoint=eye), orc.apply_nufft(x, plan),
                               rtol=1e-5, atol=1e-6)
    k = cn(rng, (2, nf, 1, 25))
    np.testing.assert_allclose(orc.apply_nufft_adj(k, plan, basis=eye), orc.apply_nufft_adj(k, plan),
                               rtol=1e-5, atol=1e-6)


# --- Toeplitz normal operator (tests/fft/test_nufft.py selfadjoint cases) ---
@pytest.mark.parametrize("ndim, ncontrasts, ncoils", list(itertools.product([2, 3], [1, 2], [1, 2])))
def test_toeplitz_delta_to_delta(ndim, ncontrasts, ncoils):
    npix = 4
    coord = _cart_coord(ndim, ncontrasts, npix)
    lead = (ncoils,) if ncontrasts == 1 else (ncoils, ncontrasts)
    img = np.zeros((*lead, *(npix,) * ndim), np.complex64)
    img[(Ellipsis,) + (npix // 2,) * ndim] = 1
    kern, kshape = orc.plan_toeplitz(coord, npix, width=12)
    np.testing.assert_allclose(orc.apply_selfadj(img, kern, kshape), img, rtol=1e-4, atol=1e-4)


def test_toeplitz_is_proportional_to_normal_operator():
    # A^H A x and the Toeplitz product agree up to deepmr's kernel
    # normalisation (a scalar): fit it and check the residual
    rng = np.random.default_rng(7)
    n = 24
    coord = np.stack([rng.uniform(-12, 12, (20, 40)), rng.uniform(-12, 12, (20, 40))], -1).astype(np.float32)
    x = cn(rng, (n, n))
    kern, kshape = orc.plan_toeplitz(coord, n, width=4)
    t = orc.apply_selfadj(x, kern, kshape).astype(np.complex128)
    plan = orc.plan_nufft(coord, n, 4, 2.0)
    aha = orc.apply_nufft_adj(orc.apply_nufft(x, plan), plan).astype(np.complex128)
    c = np.vdot(aha, t) / np.vdot(aha, aha)
    assert abs(c.imag) < 1e-2 * abs(c.real)
    assert rel(t, c * aha) < 5e-2


@pytest.mark.parametrize("ndim", [2, 3])
def test_lowrank_equals_basis_combination_of_plain(ndim):
    """A general complex subspace basis (ncoeff != nframes) against the plain,
    already pinned per-frame operators: the forward of frame f is
    sum_c basis_adjoint[f, c] * A_f(x_c) and the adjoint's coefficient c is
    sum_f basis[c, f] * A_f^H(y_f). A transposed or conjugated basis fails."""
    rng = np.random.default_rng(11 + ndim)
    nf, nc, n, m = 5, 2, (10 if ndim == 2 else 6), 21
    coord = rng.uniform(-n / 2, n / 2, (nf, 1, m, ndim)).astype(np.float32)
    plan = orc.plan_nufft(coord, n, 4, 2.0)
    ba = cn(rng, (nf, nc))
    x = cn(rng, (2, nc) + (n,) * ndim)
    y_lr = orc.apply_nufft(x, plan, basis_adjoint=ba)
    # plain: frame f sees image sum_c ba[f, c] x_c
    xf = np.einsum("fc,bc...->bf...", ba, x).astype(np.complex64)
    np.testing.assert_allclose(y_lr, orc.apply_nufft(xf, plan), rtol=2e-5, atol=2e-5)

    basis = cn(rng, (nc, nf))
    k = cn(rng, (2, nf, 1, m))
    x_lr = orc.apply_nufft_adj(k, plan, basis=basis)
    per_frame = orc.apply_nufft_adj(k, plan)  # (2, nf, *N)
    ref = np.einsum("cf,bf...->bc...", basis, per_frame)
    np.testing.assert_allclose(x_lr, ref, rtol=2e-5, atol=2e-5)


def test_lowrank_toeplitz_identity_basis_equals_per_frame():
    """Identity subspace: the coefficient pairs decouple, so both low-rank
    products (deepmr's column-sum scaling and the Gram mat-vec) equal the
    plain multi-frame Toeplitz operator."""
    rng = np.random.default_rng(21)
    nf, n = 3, 12
    coord = rng.uniform(-n / 2, n / 2, (nf, 2, 20, 2)).astype(np.float32)
    kern, ks = orc.plan_toeplitz(coord, n)
    kern_lr, ks_lr = orc.plan_toeplitz(coord, n, basis=np.eye(nf, dtype=np.complex64))
    assert ks == ks_lr and kern_lr.shape == (int(np.prod(ks)), nf, nf)
    x = cn(rng, (2, nf, n, n))
    ref = orc.apply_selfadj(x, kern, ks)
    for matvec in (False, True):
        out = orc.apply_selfadj(x, kern_lr, ks, lowrank=True, matvec=matvec)
        assert rel(out, ref) < 1e-5


def test_lowrank_toeplitz_kernel_is_basis_combination_of_frame_kernels():
    """K[v, c2, c1] = s * sum_f basis[c2, f] conj(basis[c1, f]) K_f[v] for one
    scalar s (the two normalisations): pins the (c1, c2) order and the conj."""
    rng = np.random.default_rng(22)
    nf, nc, n = 4, 2, 10
    coord = rng.uniform(-n / 2, n / 2, (nf, 2, 16, 2)).astype(np.float32)
    basis = cn(rng, (nc, nf))
    kf, ks = orc.plan_toeplitz(coord, n)  # (nf, *2N)
    klr, _ = orc.plan_toeplitz(coord, n, basis=basis)
    comb = np.einsum("if,jf,fv->vij", basis, basis.conj(), kf.reshape(nf, -1))
    s = np.vdot(comb, klr) / np.vdot(comb, comb)
    assert rel(klr, s * comb) < 1e-5
    # and the mat-vec mode is then the Gram of the low-rank NUFFT pair up to s
    x = cn(rng, (1, nc, n, n))
    y = orc.apply_selfadj(x, klr, ks, lowrank=True, matvec=True)
    per_frame = orc.apply_selfadj(np.einsum("cf,bc...->bf...", basis.conj(), x).astype(np.complex64), kf, ks)
    ref = s * np.einsum("cf,bf...->bc...", basis, per_frame)
    assert rel(y, ref) < 1e-5


@pytest.mark.parametrize("ndim, n, width", [(2, 24, 12), (2, 20, 16), (3, 12, 10)])
def test_wide_kernels_vs_ndft(ndim, n, width):
    """Widths above 8 (deepmr accepts any width): at osf 2 the KB NUFFT is
    then accurate to float32 round-off against the FP64 NDFT."""
    rng = np.random.default_rng(40 + width)
    coord = rng.uniform(-n / 2, n / 2, (120, ndim)).astype(np.float32)
    x = cn(rng, (n,) * ndim)
    y = orc.nufft(x, coord, n, width, 2.0)
    assert rel(y, ndft.ndft(x, coord, (n,) * ndim).reshape(y.shape)) < 1e-5
    plan = orc.plan_nufft(coord, n, width, 2.0)
    k = cn(rng, y.shape)
    exact = ndft.ndft_adjoint(k.reshape(-1), coord, (n,) * ndim) / np.prod(plan["os_shape"])
    assert rel(orc.apply_nufft_adj(k, plan), exact) < 1e-5
