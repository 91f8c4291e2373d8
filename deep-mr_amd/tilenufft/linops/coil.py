"""Coil sensitivity operators (drop-in for src/deepmr/linops/coil.py:11-205).

SenseOp expands a combined image into coil images, SenseAdjointOp combines
coil images with the conjugate maps (summing coils and the set axis 0).
"""

__all__ = ["SenseOp", "SenseAdjointOp"]

import numpy as np
import torch

from . import base


def _maybe_numpy(x):
    return isinstance(x, np.ndarray), torch.as_tensor(x)


class SenseOp(base.Linop):
    """Maps (nslices, ..., ny, nx) -> (nsets, nslices, ncoils, ..., ny, nx).

    sensmap: (nsets, nslices, ncoils, ny, nx) in 2D or (nsets, ncoils, nz, ny, nx)."""

    def __init__(self, ndim, sensmap, device=None, multicontrast=True):
        super().__init__(ndim)
        self.sensmap = torch.as_tensor(sensmap)
        self.device = self.sensmap.device if device is None else device
        self.sensmap = self.sensmap.to(self.device)
        self.multicontrast = multicontrast
        if multicontrast:
            self.sensmap = self.sensmap.unsqueeze(-ndim - 1)

    def forward(self, x):
        isnumpy, x = _maybe_numpy(x)
        self.sensmap = self.sensmap.to(x.device)
        x = x.unsqueeze(-self.ndim - 2 if self.multicontrast else -self.ndim - 1)
        y = self.sensmap * x
        return y.numpy(force=True) if isnumpy else y

    def _adjoint_linop(self):
        smap = self.sensmap.squeeze(-self.ndim - 1) if self.multicontrast else self.sensmap
        return SenseAdjointOp(self.ndim, smap, self.device, self.multicontrast)


class SenseAdjointOp(base.Linop):
    """Coil combination: sum over coils and sets of conj(S) * y."""

    def __init__(self, ndim, sensmap, device=None, multicontrast=True):
        super().__init__(ndim)
        self.sensmap = torch.as_tensor(sensmap)
        self.device = self.sensmap.device if device is None else device
        self.sensmap = self.sensmap.to(self.device)
        self.multicontrast = multicontrast
        if multicontrast:
            self.sensmap = self.sensmap.unsqueeze(-ndim - 1)

    def forward(self, y):
        isnumpy, y = _maybe_numpy(y)
        self.sensmap = self.sensmap.to(y.device)
        tmp = self.sensmap.conj() * y
        coil_axis = -self.ndim - 2 if self.multicontrast else -self.ndim - 1
        x = tmp.sum(dim=coil_axis).sum(dim=0)
        return x.numpy(force=True) if isnumpy else x

    def _adjoint_linop(self):
        smap = self.sensmap.squeeze(-self.ndim - 1) if self.multicontrast else self.sensmap
        return SenseOp(self.ndim, smap, self.device, self.multicontrast)
