"""ORACLE — test infrastructure only.

deepmr's autograd wiring and unrolled CG-SENSE, restated over the oracle
NUFFT, so the product's gradients can be compared with "autograd through
deepmr's reference operator" (deepmr itself cannot run here).

- ApplyNUFFT.backward calls the adjoint apply and ApplyNUFFTAdjoint.backward
  calls the forward apply (src/deepmr/fft/nufft.py:171-209, :248-286).
- E = F * C, E^H = C^H F^H, E^H E = C^H F^H F C (src/deepmr/recon/alg/linop.py:108-148).
- CGStep (src/deepmr/optim/cg.py:150-212): alpha = rs_old / <p, AHA p>,
  x += alpha p, r -= alpha AHA p, rs_new = Re<r, r>, p = r + (rs_new/rs_old) p.
"""

import numpy as np
import torch

from . import nufft64
from . import nufft_oracle as orc


class OracleNUFFT(torch.autograd.Function):
    """Forward NUFFT on the CPU oracle; backward = deepmr's adjoint (A^H / N_os)."""

    @staticmethod
    def forward(ctx, x, plan):
        ctx.plan = plan
        return torch.from_numpy(orc.apply_nufft(x.detach().cpu().numpy(), plan))

    @staticmethod
    def backward(ctx, g):
        return torch.from_numpy(orc.apply_nufft_adj(g.detach().cpu().numpy(), ctx.plan)), None


class OracleNUFFTAdjoint(torch.autograd.Function):
    """deepmr's adjoint on the CPU oracle; backward = the forward NUFFT."""

    @staticmethod
    def forward(ctx, y, plan):
        ctx.plan = plan
        return torch.from_numpy(orc.apply_nufft_adj(y.detach().cpu().numpy(), plan))

    @staticmethod
    def backward(ctx, g):
        return torch.from_numpy(orc.apply_nufft(g.detach().cpu().numpy(), ctx.plan)), None


def _dot(a, b, ndim):
    d = (a.conj() * b).reshape(*a.shape[: a.dim() - ndim], -1).sum(-1)
    return d.reshape(*d.shape, *([1] * ndim))


def cgsense(y, smaps, niter, fwd, adj, ndim=2):
    """Unrolled CG-SENSE with torch ops: smaps (C, *N), y (C, *ishape).

    Returns the image after ``niter`` CG steps from zero on E^H E x = E^H y.
    """
    def E(x):
        return fwd(smaps * x)

    def EH(k):
        return (smaps.conj() * adj(k)).sum(0)

    b = EH(y)
    r = b.clone()
    p = r
    rs_old = _dot(r, r, ndim)
    x = torch.zeros_like(b)
    for _ in range(niter):
        Ap = EH(E(p))
        alpha = rs_old / _dot(p, Ap, ndim)
        x = x + p * alpha
        r = r + Ap * (-alpha)
        rs_new = torch.real(_dot(r, r, ndim))
        p = r + p * (rs_new / rs_old)
        rs_old = rs_new
    return x


def cgsense_oracle_autograd(y, smaps, niter, plan):
    """Unrolled CG-SENSE through the oracle with deepmr's backward wiring."""
    return cgsense(y, smaps, niter, lambda v: OracleNUFFT.apply(v, plan),
                   lambda k: OracleNUFFTAdjoint.apply(k, plan))


def cgsense64(y, smaps, niter, pl64):
    """Float64 unrolled CG-SENSE (numpy), for finite differences."""
    yt = torch.from_numpy(np.asarray(y, np.complex128))
    st = torch.from_numpy(np.asarray(smaps, np.complex128))
    fwd = lambda v: torch.from_numpy(nufft64.forward(v.numpy(), pl64))  # noqa: E731
    adj = lambda k: torch.from_numpy(nufft64.adjoint(k.numpy(), pl64))  # noqa: E731
    with torch.no_grad():
        return cgsense(yt, st, niter, fwd, adj).numpy()
