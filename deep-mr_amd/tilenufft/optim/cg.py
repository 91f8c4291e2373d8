"""Conjugate gradient for the normal equations (drop-in for src/deepmr/optim/cg.py).

``CGStep`` is one differentiable CG iteration (for unrolled networks);
``cg_solve`` runs ``niter`` of them without autograd.
"""

__all__ = ["cg_solve", "CGStep"]

import numpy as np
import torch
import torch.nn as nn

from .. import linops as _linops


class CGStep(nn.Module):
    """One CG iteration on ``AHA x = AHy`` with per-batch step sizes.

    Dot products reduce over the last ``ndim`` axes (taken from ``AHA.ndim``
    when AHA is an operator), so leading axes are independent systems
    (src/deepmr/optim/cg.py:150-221).
    """

    def __init__(self, AHA, AHy, ndim=None, tol=None):
        super().__init__()
        self.ndim = getattr(AHA, "ndim", None) or ndim
        self.AHA = AHA
        self.AHy = AHy
        self.r = AHy.clone()
        self.p = self.r
        self.rsold = self.dot(self.r, self.r)
        self.rsnew = None
        self.tol = tol

    def dot(self, s1, s2):
        # sum(conj(s1) * s2) over the last ndim axes, as one reduction
        n = self.ndim
        d = torch.linalg.vecdot(s1.reshape(*s1.shape[:-n], -1), s2.reshape(*s2.shape[:-n], -1), dim=-1)
        return d.reshape(*d.shape, *([1] * n))

    def forward(self, x):
        AHAp = self.AHA(self.p)
        alpha = self.rsold / self.dot(self.p, AHAp)
        out = x + self.p * alpha
        self.r = self.r + AHAp * (-alpha)
        self.rsnew = torch.real(self.dot(self.r, self.r))
        self.p = self.r + self.p * (self.rsnew / self.rsold)
        self.rsold = self.rsnew
        return out

    def check_convergence(self):
        if self.tol is None:
            return False
        return bool((self.rsnew.sqrt() < self.tol).all())


@torch.no_grad()
def cg_solve(input, AHA, niter=10, device=None, tol=1e-4, lamda=0.0, ndim=None,
             save_history=False, verbose=False):
    """Solve AHA x = input by CG (src/deepmr/optim/cg.py:15-147). Returns (x, history)."""
    isnumpy = isinstance(input, np.ndarray)
    x = torch.as_tensor(input)
    device = x.device if device is None else device
    x = x.to(device)
    if isinstance(AHA, _linops.Linop):
        AHA = AHA.to(device)
    elif not callable(AHA):
        AHA = torch.as_tensor(AHA, dtype=x.dtype, device=device)
    AHy = x.clone()
    if lamda != 0.0:
        if isinstance(AHA, _linops.Linop):
            op = AHA + lamda * _linops.Identity(AHA.ndim)
        elif callable(AHA):
            op = lambda v, A=AHA: A(v) + lamda * v  # noqa: E731
        else:
            op = lambda v, A=AHA: A @ v + lamda * v  # noqa: E731
    else:
        op = AHA
    cg = CGStep(op, AHy, ndim, tol)
    x = 0 * x
    history = []
    out = x
    for _ in range(niter):
        out = cg(x)
        if cg.check_convergence():
            break
        x = out.clone()
        if save_history:
            r = out - AHy
            history.append(0.5 * torch.linalg.norm(r).item() ** 2
                           + lamda * torch.linalg.norm(out).item() ** 2)
    out = out.to(device)
    if isnumpy:
        out = out.numpy(force=True)
    return out, history
