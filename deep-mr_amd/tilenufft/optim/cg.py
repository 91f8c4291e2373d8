"""Conjugate gradient for the normal equations (drop-in for src/deepmr/optim/cg.py).

``CGStep`` is one differentiable CG iteration (for unrolled networks);
``cg_solve`` runs ``niter`` of them without autograd.
"""

__all__ = ["cg_solve", "CGStep"]

import math

import numpy as np
import torch
import torch.nn as nn

from .. import _core, _lib
from .. import linops as _linops


class _CGIterate(torch.autograd.Function):
    """One CG iteration's vector arithmetic on (nsys, n) complex64 device
    arrays given q = A p (tn_cg_iterate / tn_cg_iterate_backward): returns
    (x1, r1, p1, rho1); the backward returns the gradients of (x, r, p, q,
    rho) in torch's convention, so autograd carries q's gradient through A."""

    @staticmethod
    def forward(ctx, x, r, p, q, rho):
        rho = rho.to(torch.float32).contiguous()  # torch.real() of a complex tensor is a strided view
        nsys, n = r.shape
        dev = r.device
        x1, r1, p1 = torch.empty_like(r), torch.empty_like(r), torch.empty_like(r)
        rho1 = torch.empty((nsys,), dtype=torch.float32, device=dev)
        ad = torch.empty((nsys, 2), dtype=torch.complex64, device=dev)
        lib = _lib.load()
        work = torch.empty((int(lib.tn_cg_grad_work_floats(nsys, n)),), dtype=torch.float32, device=dev)
        with torch.cuda.device(dev):
            _lib.check(lib.tn_cg_iterate(_core.ptr(x), _core.ptr(r), _core.ptr(p), _core.ptr(q), _core.ptr(rho),
                                         _core.ptr(x1), _core.ptr(r1), _core.ptr(p1), _core.ptr(rho1),
                                         _core.ptr(ad), _core.ptr(work), nsys, n, _core.stream_ptr(dev)),
                       "tn_cg_iterate")
        ctx.save_for_backward(p, q, r1, rho, rho1, ad)
        ctx.set_materialize_grads(False)
        return x1, r1, p1, rho1

    @staticmethod
    def backward(ctx, gx1, gr1, gp1, grho1):
        p, q, r1, rho, rho1, ad = ctx.saved_tensors
        nsys, n = p.shape
        dev = p.device
        c = lambda g: None if g is None else g.to(torch.complex64).resolve_conj().contiguous()  # noqa: E731
        gx1, gr1, gp1 = c(gx1), c(gr1), c(gp1)
        grho1 = None if grho1 is None else grho1.to(torch.float32).contiguous()
        gr, gp, gq = torch.empty_like(p), torch.empty_like(p), torch.empty_like(p)
        grho = torch.empty((nsys,), dtype=torch.float32, device=dev)
        lib = _lib.load()
        work = torch.empty((int(lib.tn_cg_grad_work_floats(nsys, n)),), dtype=torch.float32, device=dev)
        with torch.cuda.device(dev):
            _lib.check(lib.tn_cg_iterate_backward(_core.ptr(p), _core.ptr(q), _core.ptr(r1), _core.ptr(rho),
                                                  _core.ptr(rho1), _core.ptr(ad), _core.ptr(gx1), _core.ptr(gr1),
                                                  _core.ptr(gp1), _core.ptr(grho1), _core.ptr(gr), _core.ptr(gp),
                                                  _core.ptr(gq), _core.ptr(grho), _core.ptr(work), nsys, n,
                                                  _core.stream_ptr(dev)), "tn_cg_iterate_backward")
        return gx1, gr, gp, gq, grho


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
        # complex64 systems on a HIP device take the native iteration; rs is
        # then kept real (<r, r> has an exactly zero imaginary part, so its
        # gradient is the same)
        self._native = (isinstance(AHy, torch.Tensor) and AHy.device.type == "cuda"
                        and AHy.dtype == torch.complex64 and self.ndim is not None and AHy.dim() >= self.ndim)
        if self._native:
            self._lead = tuple(AHy.shape[:AHy.dim() - self.ndim])
            self._nsys = max(1, math.prod(self._lead))
            self._rho = torch.real(self.rsold).reshape(self._nsys).to(torch.float32).contiguous()

    def dot(self, s1, s2):
        # sum(conj(s1) * s2) over the last ndim axes, as one reduction
        n = self.ndim
        d = torch.linalg.vecdot(s1.reshape(*s1.shape[:-n], -1), s2.reshape(*s2.shape[:-n], -1), dim=-1)
        return d.reshape(*d.shape, *([1] * n))

    def forward(self, x):
        if self._native and isinstance(x, torch.Tensor) and x.shape == self.r.shape and x.dtype == torch.complex64:
            AHAp = self.AHA(self.p)
            if AHAp.shape == self.p.shape and AHAp.dtype == torch.complex64:
                shape, ns = self.p.shape, self._nsys
                flat = lambda t: t.reshape(ns, -1).resolve_conj().contiguous()  # noqa: E731
                x1, r1, p1, rho1 = _CGIterate.apply(flat(x), flat(self.r), flat(self.p), flat(AHAp), self._rho)
                self.r, self.p = r1.reshape(shape), p1.reshape(shape)
                self._rho = rho1
                self.rsnew = rho1.reshape(*self._lead, *([1] * self.ndim))
                self.rsold = self.rsnew
                return x1.reshape(shape)
            return self._torch_step(x, AHAp)
        return self._torch_step(x, self.AHA(self.p))

    def _torch_step(self, x, AHAp):
        alpha = self.rsold / self.dot(self.p, AHAp)
        out = x + self.p * alpha
        self.r = self.r + AHAp * (-alpha)
        self.rsnew = torch.real(self.dot(self.r, self.r))
        self.p = self.r + self.p * (self.rsnew / self.rsold)
        self.rsold = self.rsnew
        if self._native:
            self._rho = self.rsnew.reshape(self._nsys).to(torch.float32).contiguous()
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
