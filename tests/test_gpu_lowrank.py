"""GPU parity of the low-rank (subspace) NUFFT against the oracle.

deepmr fuses a temporal subspace basis into the gather / scatter
(src/deepmr/fft/_interp/interp.py:268-357, grid.py:326-417): the forward of
frame f is sum_c basis_adjoint[f, c] * A_f(x_c), the adjoint's coefficient c
is sum_f basis[c, f] * A_f^H(y_f). Bar: relative L2 <= 1e-5 against the
oracle (complex64), the dot-product identity between the two directions, and
the autograd backward equal to the adjoint with basis = basis_adjoint^H.
"""

import numpy as np
import pytest
import torch

import tilenufft
from oracle import nufft_oracle as orc
from tilenufft import testing

pytestmark = pytest.mark.gpu

RTOL = 1e-5
DEV = "cuda:0"


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def cnormal(shape, seed):
    return testing.complex_normal(shape, seed=seed).numpy()


def spiral_frames(nframes, nviews, nsamp, n):
    return np.stack([testing.vd_spiral(nviews, nsamp, n, rotation=f * testing.GOLDEN_ANGLE)
                     for f in range(nframes)])


def koosh_frames(nframes, nspokes, nsamp, n, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-n / 2, n / 2, (nframes, nspokes, nsamp, 3)).astype(np.float32)


CASES = [
    # (ndim, N, coils, nframes, ncoeff, basis dtype, osf)
    (2, 64, 3, 12, 4, "complex", 2.0),
    (2, 48, 2, 5, 5, "real", 1.25),   # ncoeff == nframes (deepmr's own case)
    (3, 16, 2, 6, 3, "complex", 1.5),
]


@pytest.mark.parametrize("grid_mode", [0, 3, 4])
@pytest.mark.parametrize("ndim, n, ncoil, nframes, ncoeff, kind, osf", CASES)
def test_lowrank_nufft_matches_oracle(ndim, n, ncoil, nframes, ncoeff, kind, osf, grid_mode):
    if grid_mode == 4 and ndim != 3:
        pytest.skip("register-tile gridding is 3D")
    print(f"low-rank {ndim}D N={n} coils={ncoil} frames={nframes} coeff={ncoeff} {kind} grid_mode={grid_mode}",
          flush=True)
    if ndim == 2:
        coord = spiral_frames(nframes, 2, 256, n)
    else:
        coord = koosh_frames(nframes, 4, 40, n, 3)
    rng = np.random.default_rng(ndim * 100 + ncoeff)
    if kind == "complex":
        ba = (rng.standard_normal((nframes, ncoeff)) + 1j * rng.standard_normal((nframes, ncoeff))).astype(np.complex64)
    else:
        ba = rng.standard_normal((nframes, ncoeff)).astype(np.float32)
    basis = np.ascontiguousarray(ba.conj().T)
    x = cnormal((ncoil, ncoeff) + (n,) * ndim, 5)
    plan = tilenufft.fft.plan_nufft(coord, n, 4, osf, device=DEV, grid_mode=grid_mode)
    y = tilenufft.fft.apply_nufft(torch.from_numpy(x).to(DEV), plan,
                                  basis_adjoint=torch.from_numpy(ba)).cpu().numpy()
    k = cnormal(y.shape, 6)
    xa = tilenufft.fft.apply_nufft_adj(torch.from_numpy(k).to(DEV), plan,
                                       basis=torch.from_numpy(basis)).cpu().numpy()
    p_ref = orc.plan_nufft(coord, n, 4, osf)
    y_ref = orc.apply_nufft(x, p_ref, basis_adjoint=ba.astype(np.complex64))
    xa_ref = orc.apply_nufft_adj(k, p_ref, basis=basis.astype(np.complex64))
    assert y.shape == y_ref.shape == (ncoil, nframes) + coord.shape[1:-1]
    assert xa.shape == xa_ref.shape == x.shape
    assert rel(y, y_ref) < RTOL
    assert rel(xa, xa_ref) < RTOL
    # adjoint pair: <A x, k> = prod(os_shape) <x, nufft_adj(k)> with basis = ba^H
    nos = float(np.prod(plan.os_shape))
    lhs = np.vdot(y.astype(np.complex128), k.astype(np.complex128))  # sums in FP64
    rhs = nos * np.vdot(x.astype(np.complex128), xa.astype(np.complex128))
    # on the Cauchy-Schwarz scale |y| |k|: for random k, <y, k> itself is only
    # about |y| |k| / sqrt(len(k)), which would magnify FP32 rounding; a wrong
    # conjugate or basis shows up at ~1e-2 on this scale
    assert abs(lhs - rhs) / (np.linalg.norm(y) * np.linalg.norm(k)) < 1e-6


def test_lowrank_weight_fused_matches_oracle():
    n, nframes, ncoeff = 32, 8, 3
    print("low-rank weight fusion", flush=True)
    coord = spiral_frames(nframes, 2, 128, n)
    rng = np.random.default_rng(2)
    ba = (rng.standard_normal((nframes, ncoeff)) + 1j * rng.standard_normal((nframes, ncoeff))).astype(np.complex64)
    w = rng.uniform(0.2, 1.0, coord.shape[:-1]).astype(np.float32)
    x = cnormal((2, ncoeff, n, n), 8)
    plan = tilenufft.fft.plan_nufft(coord, n, 4, 2.0, device=DEV)
    y = tilenufft.fft.apply_nufft(torch.from_numpy(x).to(DEV), plan, basis_adjoint=ba,
                                  weight=torch.from_numpy(w)).cpu().numpy()
    p_ref = orc.plan_nufft(coord, n, 4, 2.0)
    assert rel(y, orc.apply_nufft(x, p_ref, weight=w, basis_adjoint=ba)) < RTOL
    k = cnormal(y.shape, 9)
    xa = tilenufft.fft.apply_nufft_adj(torch.from_numpy(k).to(DEV), plan, basis=ba.conj().T,
                                       weight=torch.from_numpy(w)).cpu().numpy()
    assert rel(xa, orc.apply_nufft_adj(k, p_ref, weight=w, basis=np.ascontiguousarray(ba.conj().T))) < RTOL


def test_lowrank_interpolator_layer_matches_oracle():
    """apply_interpolation / apply_gridding with a basis (the _interp layer)."""
    n, nframes, ncoeff = 40, 6, 2
    print("low-rank interpolator layer", flush=True)
    rng = np.random.default_rng(4)
    coord = rng.uniform(0, n, (nframes, 3, 50, 2)).astype(np.float32)
    ba = (rng.standard_normal((nframes, ncoeff)) + 1j * rng.standard_normal((nframes, ncoeff))).astype(np.complex64)
    interp = tilenufft.fft._interp.plan_interpolator(coord, n, 4, 3.0, device=DEV)
    ref = orc.plan_interpolator(coord, n, 4, 3.0)
    g = cnormal((3, ncoeff, n, n), 10)
    y = tilenufft.fft._interp.apply_interpolation(torch.from_numpy(g).to(DEV), interp, ba).cpu().numpy()
    assert rel(y, orc.apply_interpolation(g, ref, ba)) < RTOL
    k = cnormal(y.shape, 11)
    bs = np.ascontiguousarray(ba.conj().T)
    out = tilenufft.fft._interp.apply_gridding(torch.from_numpy(k).to(DEV), interp, bs).cpu().numpy()
    assert rel(out, orc.apply_gridding(k, ref, bs)) < RTOL


def test_lowrank_autograd_backward_is_adjoint():
    n, nframes, ncoeff = 32, 6, 2
    print("low-rank autograd", flush=True)
    coord = spiral_frames(nframes, 2, 128, n)
    rng = np.random.default_rng(3)
    ba = torch.from_numpy((rng.standard_normal((nframes, ncoeff))
                           + 1j * rng.standard_normal((nframes, ncoeff))).astype(np.complex64))
    plan = tilenufft.fft.plan_nufft(coord, n, 4, 2.0, device=DEV)
    x = torch.from_numpy(cnormal((2, ncoeff, n, n), 12)).to(DEV).requires_grad_(True)
    y = tilenufft.fft.apply_nufft(x, plan, basis_adjoint=ba)
    g = torch.from_numpy(cnormal(tuple(y.shape), 13)).to(DEV)
    (y.conj() * g).real.sum().backward()
    # d/dx Re<y, g> with y = A x is A^H g (exact VJP); A^H = nos * deepmr adjoint
    nos = float(np.prod(plan.os_shape))
    ref = nos * tilenufft.fft.apply_nufft_adj(g, plan, basis=ba.conj().t())
    assert rel(x.grad.cpu().numpy(), ref.cpu().numpy()) < 1e-5
