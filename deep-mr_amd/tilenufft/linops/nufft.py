"""NUFFT linear operators (drop-in for src/deepmr/linops/nufft.py:12-186).

The density weight is stored as ``weight ** 0.5`` and applied in both
directions (fused into the native kernels when it has the trajectory's
shape); ``.H`` shares the device plan.
"""

__all__ = ["NUFFTOp", "NUFFTAdjointOp", "NUFFTGramOp"]

import torch

from .. import fft as _fft
from . import base


class NUFFTOp(base.Linop):
    """Forward NUFFT. Images (nslices, nsets, ncoils, ncontrasts, ny, nx) or
    (nsets, ncoils, ncontrasts, nz, ny, nx) -> k-space (..., nviews, nsamples)."""

    def __init__(self, coord=None, shape=None, basis_adjoint=None, weight=None, device="cpu",
                 threadsperblock=128, width=4, oversamp=1.25):
        if coord is not None and shape is not None:
            super().__init__(ndim=coord.shape[-1])
            self.nufft_plan = _fft.plan_nufft(coord, shape, width, oversamp, device)
        else:
            super().__init__(ndim=None)
            self.nufft_plan = None
        self.weight = None if weight is None else torch.as_tensor(weight ** 0.5, device=_dev(self, device))
        self.basis_adjoint = (None if basis_adjoint is None
                              else torch.as_tensor(basis_adjoint, device=_dev(self, device)))
        self.threadsperblock = threadsperblock

    def forward(self, x):
        return _fft.apply_nufft(x, self.nufft_plan, self.basis_adjoint, self.weight,
                                threadsperblock=self.threadsperblock)

    def _adjoint_linop(self):
        basis = None if self.basis_adjoint is None else self.basis_adjoint.conj().T
        weight = None if self.weight is None else self.weight ** 2
        op = NUFFTAdjointOp(basis=basis, weight=weight, threadsperblock=self.threadsperblock)
        op.ndim = self.ndim
        op.nufft_plan = self.nufft_plan
        return op


class NUFFTAdjointOp(base.Linop):
    """deepmr's adjoint NUFFT (A^H / prod(os_shape)) as an operator."""

    def __init__(self, coord=None, shape=None, basis=None, weight=None, device="cpu",
                 threadsperblock=128, width=4, oversamp=1.25, **kwargs):
        if coord is not None and shape is not None:
            super().__init__(ndim=coord.shape[-1])
            self.nufft_plan = _fft.plan_nufft(coord, shape, width, oversamp, device)
        else:
            super().__init__(ndim=None)
            self.nufft_plan = None
        self.weight = None if weight is None else torch.as_tensor(weight ** 0.5, device=_dev(self, device))
        self.basis = None if basis is None else torch.as_tensor(basis, device=_dev(self, device))
        self.threadsperblock = threadsperblock

    def forward(self, y):
        return _fft.apply_nufft_adj(y, self.nufft_plan, self.basis, self.weight,
                                    threadsperblock=self.threadsperblock)

    def _adjoint_linop(self):
        basis_adjoint = None if self.basis is None else self.basis.conj().T
        weight = None if self.weight is None else self.weight ** 2
        op = NUFFTOp(basis_adjoint=basis_adjoint, weight=weight, threadsperblock=self.threadsperblock)
        op.ndim = self.ndim
        op.nufft_plan = self.nufft_plan
        return op


def _dev(op, device):
    plan = getattr(op, "nufft_plan", None)
    if plan is not None:
        return plan.interpolator.device
    return None if device is None or torch.device(device).type == "cpu" else device


class NUFFTGramOp(base.Linop):
    """Self-adjoint NUFFT (A^H A) through a Toeplitz kernel
    (src/deepmr/linops/nufft.py:189-241)."""

    def __init__(self, coord, shape, basis=None, weight=None, device="cpu", threadsperblock=128,
                 width=6, **kwargs):
        super().__init__(ndim=coord.shape[-1])
        self.toeplitz_kern = _fft.plan_toeplitz_nufft(coord, shape, basis, weight, width, device)
        self.threadsperblock = threadsperblock

    def forward(self, x):
        return _fft.apply_nufft_selfadj(x, self.toeplitz_kern, threadsperblock=self.threadsperblock)

    def _adjoint_linop(self):
        return self
