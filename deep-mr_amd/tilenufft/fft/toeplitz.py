"""Toeplitz (self-adjoint) NUFFT: A^H A as an FFT convolution.

Drop-in for deepmr's ``plan_toeplitz_nufft`` / ``apply_nufft_selfadj``
(src/deepmr/fft/toeplitz.py:15-157, src/deepmr/fft/nufft.py:135-168, :325-375,
:609-669, src/deepmr/fft/_interp/toeplitz.py). The point-spread kernel is
computed once with the hand-written adjoint NUFFT on a 2x grid (the low-rank
gridding kernel when a subspace basis is given); each application is
zero-pad -> FFT -> per-voxel product -> IFFT -> crop (rocFFT via torch).

Low-rank product. deepmr's per-voxel routine (_interp/toeplitz.py:109-130,
``out[j] += K[i][j] * in[j]``) scales each coefficient by the column sum of
the voxel's (ncoeff x ncoeff) kernel; that is reproduced by default
(``lowrank_product="deepmr"``). ``lowrank_product="matvec"`` forms the Gram
matrix product out[i] = sum_j K[i][j] in[j], which is A^H A of the subspace
NUFFT up to the kernel normalisation.
"""

__all__ = ["GramMatrix", "plan_toeplitz_nufft", "apply_nufft_selfadj", "ApplyNUFFTSelfAdjoint"]

from dataclasses import dataclass

import numpy as np
import torch

from .. import _core
from .._signal.resize import resize


@dataclass
class GramMatrix:
    """Toeplitz kernel (mirror of deepmr's GramMatrix, toeplitz.py:165-196)."""

    value: torch.Tensor
    shape: tuple
    ndim: int
    device: str
    islowrank: bool
    lowrank_product: str = "deepmr"
    _colsum: torch.Tensor = None

    def to(self, device):
        if torch.device(device).type != "cpu":
            dev = _core.compute_device(device)
            self.value = self.value.to(dev)
            self._colsum = None
            self.device = str(dev)
        return self

    def colsum(self):
        """(voxels, ncoeff) column sums of the low-rank kernel (deepmr's product)."""
        if self._colsum is None or self._colsum.device != self.value.device:
            self._colsum = self.value.sum(dim=1).contiguous()
        return self._colsum


def plan_toeplitz_nufft(coord, shape, basis=None, dcf=None, width=4, device="cpu",
                        lowrank_product="deepmr"):
    """Spatio-temporal PSF kernel for A^H A (src/deepmr/fft/toeplitz.py:15-157).

    coord (ncontrasts, nviews, nsamples, ndim) in [-N/2, N/2); ``shape`` the
    image size. The kernel lives on a grid oversampled by 2 (1 on Cartesian
    axes), computed with nufft_adj at deepmr's default oversampling (1.25).
    ``basis`` (ncoeff, ncontrasts) gives the subspace kernel, stored as
    (voxels, ncoeff, ncoeff) like deepmr's; ``lowrank_product`` selects how it
    is applied ("deepmr" or "matvec").
    """
    from . import nufft_adj

    if lowrank_product not in ("deepmr", "matvec"):
        raise ValueError("lowrank_product must be 'deepmr' or 'matvec'")
    coord = torch.as_tensor(coord, dtype=torch.float32).cpu()
    ndim = coord.shape[-1]
    oversamp = np.asarray([2.0] * ndim)
    shape = np.asarray([shape] * ndim if np.isscalar(shape) else list(shape)[-ndim:], dtype=np.int16)
    dev = _core.compute_device(device)
    if dcf is None:
        dcf = torch.ones(coord.shape[:-1], dtype=torch.float32)
    dcf = torch.as_tensor(dcf, dtype=torch.float32).cpu()
    isreal = False
    if basis is not None:
        basis = torch.as_tensor(basis).cpu()
        isreal = not torch.is_complex(basis)
        basis = basis.to(torch.float32 if isreal else torch.complex64).resolve_conj()
        ncoeff = basis.shape[0]
        adj = basis.conj().T  # (ncontrasts, ncoeff)
        # delta[c1, f, ...] = conj(basis[c1, f])  (toeplitz.py:97-101)
        delta = torch.ones((ncoeff, *coord.shape[:-1]), dtype=adj.dtype)
        delta = (delta.permute(*range(delta.dim() - 1, -1, -1)) * adj).permute(
            *range(delta.dim() - 1, -1, -1))
    else:
        delta = torch.ones(coord.shape[:-1], dtype=torch.float32)
    is_cart = _core.cartesian_axes(coord.numpy(), shape)
    oversamp[is_cart] = 1.0
    kshape = np.ceil(oversamp * shape).astype(np.int16)
    # deepmr multiplies coord (x, y, z) by oversamp in (z, y, x) order
    kcoord = coord * torch.as_tensor(oversamp, dtype=torch.float32)
    psf = nufft_adj((dcf * delta).to(dev), kcoord, [int(s) for s in kshape], basis, dev, width=width)
    ax = tuple(range(-ndim, 0))
    kern = torch.fft.fftshift(torch.fft.fftn(torch.fft.ifftshift(psf, dim=ax), dim=ax, norm="ortho"), dim=ax)
    if kern.shape[0] == 1:
        kern = kern[0]
    if isreal:
        kern = kern.real
    kern = torch.fft.ifftshift(kern, dim=ax)
    if basis is not None:  # (c1, c2, voxels) -> (voxels, c2, c1)  (toeplitz.py:143-147)
        kern = kern.reshape(*kern.shape[:2], -1).permute(2, 1, 0).contiguous()
    kern = kern / torch.mean(torch.abs(kern[kern != 0]))
    kern = torch.nan_to_num(kern)
    return GramMatrix(kern, tuple(int(s) for s in kshape), ndim, str(dev), basis is not None,
                      lowrank_product)


def _kspace_product(k, kern, ndim):
    """Per-voxel kernel product in k-space. Low rank: k-space (..., ncoeff,
    *G) as (batch, ncoeff, voxels) against the (voxels, ncoeff, ncoeff)
    kernel (_interp/toeplitz.py:51-84)."""
    if not kern.islowrank:
        return kern.value * k
    shp = k.shape
    kb = k.reshape(-1, shp[-ndim - 1], int(np.prod(shp[-ndim:])))
    if kern.lowrank_product == "deepmr":
        return (kb * kern.colsum().T.unsqueeze(0)).reshape(shp)
    val = kern.value.permute(1, 2, 0)  # (i, j, voxels)
    return (val.unsqueeze(0) * kb.unsqueeze(1)).sum(dim=2).reshape(shp)


def _selfadj(image, kern):
    isnumpy = isinstance(image, np.ndarray)
    x = torch.as_tensor(image)
    x = x.to(torch.float32) if x.dtype in (torch.float16, torch.float32, torch.float64) else x.to(torch.complex64)
    ndim = kern.ndim
    oshape = tuple(x.shape[-ndim:])
    odevice = x.device
    dev = torch.device(kern.device)
    x = x.to(dev)
    x = resize(x, list(x.shape[:-ndim]) + list(kern.shape))
    ax = tuple(range(-ndim, 0))
    k = torch.fft.fftn(x, dim=ax, norm="ortho")
    y = torch.fft.ifftn(_kspace_product(k, kern, ndim), dim=ax, norm="ortho")
    y = resize(y, list(y.shape[:-ndim]) + list(oshape)).to(odevice)
    return y.numpy(force=True) if isnumpy else y


class ApplyNUFFTSelfAdjoint(torch.autograd.Function):
    """A^H A via the Toeplitz kernel; self-adjoint, so it is its own backward."""

    @staticmethod
    def forward(image, toeplitz_kern, device, threadsperblock):
        return _selfadj(image, toeplitz_kern)

    @staticmethod
    def setup_context(ctx, inputs, output):
        ctx.toeplitz_kern = inputs[1]
        ctx.set_materialize_grads(False)

    @staticmethod
    def backward(ctx, grad):
        if grad is None:
            return None, None, None, None
        return _selfadj(grad, ctx.toeplitz_kern), None, None, None


def apply_nufft_selfadj(image, toeplitz_kern, device=None, threadsperblock=128):
    """Apply A^H A through the Toeplitz kernel (src/deepmr/fft/nufft.py:351-375, :609-669)."""
    if device is not None:
        toeplitz_kern.to(device)
    return ApplyNUFFTSelfAdjoint.apply(image, toeplitz_kern, device, threadsperblock)
