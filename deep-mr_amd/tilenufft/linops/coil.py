"""Coil sensitivity operators (drop-in for src/deepmr/linops/coil.py:11-205).

SenseOp expands a combined image into coil images, SenseAdjointOp combines
coil images with the conjugate maps (summing coils and the set axis 0).
"""

__all__ = ["SenseOp", "SenseAdjointOp"]

import math

import numpy as np
import torch

from . import base
from .. import _core, _lib


def _maybe_numpy(x):
    return isinstance(x, np.ndarray), torch.as_tensor(x)


def _combine(smaps, y, nsets, nb, ncoils, nk, npix):
    """Native coil combination: smaps (nsets, nb, ncoils, npix), y (nsets, nb,
    ncoils, nk, npix) contiguous complex64 -> (nb, nk, npix)."""
    x = torch.empty((nb, nk, npix), dtype=torch.complex64, device=y.device)
    with torch.cuda.device(y.device):
        _lib.check(_lib.load().tn_coil_combine(_core.ptr(smaps), _core.ptr(y), _core.ptr(x), nsets, nb, ncoils,
                                               nk, npix, _core.stream_ptr(y.device)), "tn_coil_combine")
    return x


class _SenseCombine(torch.autograd.Function):
    """x = sum over sets and coils of conj(S) * y (S constant); backward
    grad_y = S * grad_x."""

    @staticmethod
    def forward(ctx, smaps, y):
        nsets, nb, nc, nk, npix = y.shape
        ctx.save_for_backward(smaps)
        return _combine(smaps, y.contiguous(), nsets, nb, nc, nk, npix)

    @staticmethod
    def backward(ctx, gx):
        (smaps,) = ctx.saved_tensors
        return None, smaps[:, :, :, None, :] * gx[None, :, None, :, :]


class _SenseExpand(torch.autograd.Function):
    """y = S * x (S constant); backward grad_x = sum over sets and coils of
    conj(S) * grad_y (the native combination instead of a strided sum)."""

    @staticmethod
    def forward(ctx, smaps, x):
        ctx.save_for_backward(smaps)
        return smaps[:, :, :, None, :] * x[None, :, None, :, :]

    @staticmethod
    def backward(ctx, gy):
        (smaps,) = ctx.saved_tensors
        nsets, nb, nc, nk, npix = gy.shape
        return None, _combine(smaps, gy.to(torch.complex64).contiguous(), nsets, nb, nc, nk, npix)


def _native_layout(sensmap, other, ndim, multicontrast):
    """Shapes of the native path, or None where the torch expression must be
    used (maps that need gradients, other devices or dtypes, no set axis).
    Returns (full, coil_pos, nsets, nb, ncoils, nk, npix, smaps (nsets, nb,
    ncoils, npix) contiguous)."""
    if sensmap.requires_grad or sensmap.device.type != "cuda" or other.device != sensmap.device:
        return None
    if sensmap.dtype != torch.complex64 or other.dtype not in (torch.complex64, torch.float32):
        return None
    try:
        full = list(torch.broadcast_shapes(sensmap.shape, other.shape))
    except RuntimeError:
        return None
    coil_pos = len(full) - ndim - (2 if multicontrast else 1)
    if coil_pos < 1:
        return None
    nsets, ncoils = full[0], full[coil_pos]
    nb = math.prod(full[1:coil_pos])
    nk = full[-ndim - 1] if multicontrast else 1
    npix = math.prod(full[-ndim:])
    smap_shape = list(full)
    if multicontrast:
        smap_shape[-ndim - 1] = 1
    smaps = sensmap.expand(smap_shape).contiguous().reshape(nsets, nb, ncoils, npix)
    return full, coil_pos, nsets, nb, ncoils, nk, npix, smaps


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
        lay = _native_layout(self.sensmap, x, self.ndim, self.multicontrast)
        if lay is not None and (x.dim() < len(lay[0]) or x.shape[0] == 1):
            full, coil_pos, nsets, nb, nc, nk, npix, smaps = lay
            xs = [1] + full[1:coil_pos] + [1] + full[coil_pos + 1:]
            x3 = torch.broadcast_to(x.to(torch.complex64), xs).reshape(nb, nk, npix)
            y = _SenseExpand.apply(smaps, x3).reshape(full)
        else:
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
        lay = _native_layout(self.sensmap, y, self.ndim, self.multicontrast)
        if lay is not None:
            full, coil_pos, nsets, nb, nc, nk, npix, smaps = lay
            y5 = torch.broadcast_to(y.to(torch.complex64), full).reshape(nsets, nb, nc, nk, npix)
            out_shape = full[1:coil_pos] + ([nk] if self.multicontrast else []) + full[-self.ndim:]
            x = _SenseCombine.apply(smaps, y5).reshape(out_shape)
        else:
            tmp = self.sensmap.conj() * y
            coil_axis = -self.ndim - 2 if self.multicontrast else -self.ndim - 1
            x = tmp.sum(dim=coil_axis).sum(dim=0)
        return x.numpy(force=True) if isnumpy else x

    def _adjoint_linop(self):
        smap = self.sensmap.squeeze(-self.ndim - 1) if self.multicontrast else self.sensmap
        return SenseOp(self.ndim, smap, self.device, self.multicontrast)
