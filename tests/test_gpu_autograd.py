"""Autograd through the hand-written kernels (GPU).

- exact mode: directional derivatives of losses through the forward, the
  adjoint and an unrolled 10-iteration CG-SENSE (64x64, 4 coils, 32 spokes)
  against float64 central finite differences of the FP64 oracle;
- deepmr mode: gradients equal autograd through deepmr's wiring restated on
  the CPU oracle (oracle/autograd_ref.py); exact mode differs from it only by
  the prod(os_shape) factor on k-space-side gradients.
"""

import numpy as np
import pytest
import torch

import tilenufft
from oracle import autograd_ref, nufft64
from oracle import nufft_oracle as orc
from tilenufft import testing

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


@pytest.fixture(autouse=True)
def _exact_mode():
    tilenufft.fft.set_autograd_mode("exact")
    yield
    tilenufft.fft.set_autograd_mode("exact")


def cn(shape, seed, dtype=np.complex128):
    return testing.complex_normal(shape, seed=seed).numpy().astype(dtype)


def fd(fun, x, v, eps):
    return (fun(x + eps * v) - fun(x - eps * v)) / (2 * eps)


def test_forward_and_adjoint_vjp_vs_fp64_fd():
    n = 32
    print("VJP of forward/adjoint vs FP64 finite differences", flush=True)
    coord = testing.golden_angle_radial(16, 64, n)
    pl64 = nufft64.plan(coord, n, 4, 1.5)
    plan = tilenufft.fft.plan_nufft(coord, n, 4, 1.5, device=DEV)
    x0, v = cn((2, n, n), 1), cn((2, n, n), 2)
    b = cn((2, 16, 64), 3)

    # L(x) = ||A x - b||^2
    xt = torch.from_numpy(x0.astype(np.complex64)).to(DEV).requires_grad_(True)
    loss = (tilenufft.fft.apply_nufft(xt, plan) - torch.from_numpy(b.astype(np.complex64)).to(DEV)).abs().pow(2).sum()
    loss.backward()
    g = xt.grad.cpu().numpy().astype(np.complex128)
    d_fd = fd(lambda z: np.sum(np.abs(nufft64.forward(z, pl64) - b) ** 2), x0, v, 1e-4)
    d_an = np.real(np.vdot(g, v))
    assert abs(d_an - d_fd) / abs(d_fd) < 1e-4

    # L(y) = ||B y - c||^2 with B = deepmr's adjoint
    y0, w = cn((2, 16, 64), 4), cn((2, 16, 64), 5)
    c = cn((2, n, n), 6) * 1e-4
    yt = torch.from_numpy(y0.astype(np.complex64)).to(DEV).requires_grad_(True)
    loss = (tilenufft.fft.apply_nufft_adj(yt, plan) - torch.from_numpy(c.astype(np.complex64)).to(DEV)).abs().pow(2).sum()
    loss.backward()
    g = yt.grad.cpu().numpy().astype(np.complex128)
    d_fd = fd(lambda z: np.sum(np.abs(nufft64.adjoint(z, pl64) - c) ** 2), y0, w, 1e-3)
    assert abs(np.real(np.vdot(g, w)) - d_fd) / abs(d_fd) < 1e-4


def _cgsense_problem():
    n, ncoil, nspokes, nsamp = 64, 4, 32, 128
    coord = testing.golden_angle_radial(nspokes, nsamp, n)
    smaps = testing.gaussian_coil_maps(ncoil, (n, n), seed=1)
    truth = cn((n, n), 7)
    pl64 = nufft64.plan(coord, n, 4, 1.25)
    y = nufft64.forward(smaps * truth, pl64)
    target = cn((n, n), 8) * 0.1
    return n, coord, smaps, y, target, pl64


def test_unrolled_cgsense_grad_vs_fp64_fd():
    print("unrolled CG-SENSE 64x64 / 4 coils / 32 spokes, 10 iters: grad vs FP64 FD", flush=True)
    n, coord, smaps, y, target, pl64 = _cgsense_problem()
    model = tilenufft.recon.UnrolledCGSENSE(torch.from_numpy(coord), (n, n),
                                            torch.from_numpy(smaps).to(DEV), niter=10, device=DEV)
    yt = torch.from_numpy(y.astype(np.complex64)).to(DEV).requires_grad_(True)
    tt = torch.from_numpy(target.astype(np.complex64)).to(DEV)
    x = model(yt)
    assert x.shape == (n, n)
    loss = (x - tt).abs().pow(2).mean()
    loss.backward()
    g = yt.grad.cpu().numpy().astype(np.complex128)

    v = cn(y.shape, 9)
    v *= np.linalg.norm(y) / np.linalg.norm(v)

    def L(z):
        xz = autograd_ref.cgsense64(z, smaps, 10, pl64)
        return np.mean(np.abs(xz - target) ** 2)

    d_fd = fd(L, y, v, 1e-5)
    d_an = np.real(np.vdot(g, v))
    assert abs(d_an - d_fd) / abs(d_fd) < 2e-3
    # the forward recon itself matches the FP64 restatement
    x64 = autograd_ref.cgsense64(y, smaps, 10, pl64)
    assert np.linalg.norm(x.detach().cpu().numpy() - x64) / np.linalg.norm(x64) < 1e-3


def test_unrolled_cgsense_grad_vs_deepmr_wiring():
    print("unrolled CG-SENSE grad vs autograd through deepmr's wiring (oracle)", flush=True)
    n, coord, smaps, y, target, _ = _cgsense_problem()
    niter = 3
    plan_ref = orc.plan_nufft(coord, n, 4, 1.25)
    y32 = y.astype(np.complex64)
    s_ref = torch.from_numpy(smaps)
    yr = torch.from_numpy(y32).requires_grad_(True)
    xr = autograd_ref.cgsense_oracle_autograd(yr, s_ref, niter, plan_ref)
    ((xr - torch.from_numpy(target.astype(np.complex64))).abs().pow(2).mean()).backward()
    g_ref = yr.grad.numpy()

    grads = {}
    for mode in ("deepmr", "exact"):
        tilenufft.fft.set_autograd_mode(mode)
        model = tilenufft.recon.UnrolledCGSENSE(torch.from_numpy(coord), (n, n),
                                                torch.from_numpy(smaps).to(DEV), niter=niter, device=DEV)
        yt = torch.from_numpy(y32).to(DEV).requires_grad_(True)
        x = model(yt)
        ((x - torch.from_numpy(target.astype(np.complex64)).to(DEV)).abs().pow(2).mean()).backward()
        grads[mode] = yt.grad.cpu().numpy()
        assert np.linalg.norm(x.detach().cpu().numpy() - xr.detach().numpy()) / np.linalg.norm(xr.detach().numpy()) < 1e-4
    nos = float(np.prod(plan_ref["os_shape"]))
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)  # noqa: E731
    assert rel(grads["deepmr"], g_ref) < 1e-3
    assert rel(grads["exact"] * nos, g_ref) < 1e-3


def test_unrolled_cgsense_graph_replays_equal_eager():
    """capture_unrolled: the HIP-graph forward and backward of the unrolled
    CG-SENSE give the eager module's image and k-space gradient on every
    replay. The replays take different k-space, and eager NUFFT work and
    host synchronisations run between them."""
    print("unrolled CG-SENSE graph capture: replays vs eager", flush=True)
    n, coord, smaps, y, target, _ = _cgsense_problem()
    tt = torch.from_numpy(target.astype(np.complex64)).to(DEV)

    def make():
        return tilenufft.recon.UnrolledCGSENSE(torch.from_numpy(coord), (n, n),
                                               torch.from_numpy(smaps).to(DEV), niter=10, device=DEV)

    def run(model, yv):
        yy = yv.detach().clone().requires_grad_(True)
        x = model(yy)
        (x - tt).abs().pow(2).mean().backward()
        return x.detach().clone(), yy.grad.detach().clone()

    ys = [torch.from_numpy((y * s + cn(y.shape, 20 + k) * 0.05).astype(np.complex64)).to(DEV)
          for k, s in enumerate((1.0, 0.5, -2.0))]
    eager = make()
    refs = [run(eager, yv) for yv in ys]
    graphed = tilenufft.recon.capture_unrolled(make(), ys[0].clone().requires_grad_(True))
    side_plan = tilenufft.fft.plan_nufft(coord, n, 4, 2.0, device=DEV)
    side_img = torch.from_numpy(cn((2, n, n), 30).astype(np.complex64)).to(DEV)
    for rep in range(2):
        for k, (yv, (x_ref, g_ref)) in enumerate(zip(ys, refs)):
            x, g = run(graphed, yv)
            tilenufft.fft.apply_nufft_adj(tilenufft.fft.apply_nufft(side_img, side_plan), side_plan)
            torch.cuda.synchronize()
            ex = float((x - x_ref).abs().norm() / x_ref.abs().norm())
            eg = float((g - g_ref).abs().norm() / g_ref.abs().norm())
            assert ex < 1e-6 and eg < 1e-6, (rep, k, ex, eg)
