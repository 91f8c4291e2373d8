"""Centred FFT helpers (torch.fft / rocFFT), mirroring src/deepmr/fft/fft.py:9-175.

These are the library FFT around the NUFFT; the NUFFT path itself does not
call them (its centring shifts are folded into the native kernels).
"""

__all__ = ["fft", "ifft"]

import numpy as np
import torch


def _axes(axes, ndim):
    if axes is None:
        return tuple(range(ndim))
    return tuple(a % ndim for a in sorted(axes))


def fft(input, axes=None, norm="ortho", centered=True):
    """Centred n-dimensional FFT: fftshift(fftn(ifftshift(x)))."""
    isnumpy = isinstance(input, np.ndarray)
    x = torch.as_tensor(input)
    ax = _axes(axes, x.ndim)
    if centered:
        y = torch.fft.fftshift(torch.fft.fftn(torch.fft.ifftshift(x, dim=ax), dim=ax, norm=norm), dim=ax)
    else:
        y = torch.fft.fftn(x, dim=ax, norm=norm)
    return np.asarray(y) if isnumpy else y


def ifft(input, axes=None, norm="ortho", centered=True):
    """Centred n-dimensional inverse FFT: fftshift(ifftn(ifftshift(x)))."""
    isnumpy = isinstance(input, np.ndarray)
    x = torch.as_tensor(input)
    ax = _axes(axes, x.ndim)
    if centered:
        y = torch.fft.fftshift(torch.fft.ifftn(torch.fft.ifftshift(x, dim=ax), dim=ax, norm=norm), dim=ax)
    else:
        y = torch.fft.ifftn(x, dim=ax, norm=norm)
    return np.asarray(y) if isnumpy else y
