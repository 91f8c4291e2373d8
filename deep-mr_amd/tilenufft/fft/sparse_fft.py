"""Cartesian sparse FFT on MI355X (drop-in for deepmr.fft.sparse_fft).

Mirrors src/deepmr/fft/sparse_fft.py:21-503 and src/deepmr/fft/_sparse/*:
``prepare_sampling``, ``apply_sparse_fft``, ``apply_sparse_ifft``,
``plan_toeplitz_fft``, ``apply_sparse_fft_selfadj`` and the autograd
Functions, with deepmr's layouts, dtype casts and numpy-in / numpy-out rule.

An index i (x first) on the centred grid is the frequency i - N//2, so the
sparse FFT is the NUFFT chain with width-1 taps, no oversampling and no
apodization:

  forward  tn_apod_pad (the folded ifftshift) -> rocFFT -> tn_interp gather
           at (i + N - N//2) mod N (the folded fftshift)
  adjoint  tn_grid scatter-add (duplicates accumulate, deterministic,
           every cell written) -> inverse rocFFT -> tn_crop_apod

Subspace bases use the low-rank kernels. 1D masks run on the 2D kernels with
a unit y axis.
"""

__all__ = [
    "Sampling",
    "prepare_sampling",
    "apply_sparse_fft",
    "apply_sparse_ifft",
    "plan_toeplitz_fft",
    "apply_sparse_fft_selfadj",
    "SparseGramMatrix",
    "ApplySparseFFT",
    "ApplySparseIFFT",
    "ApplySparseFFTSelfadjoint",
]

from dataclasses import dataclass, field

import numpy as np
import torch

from .. import _core
from . import nufft as _nufft
from .toeplitz import _kspace_product


@dataclass
class Sampling:
    """Sampling mask (mirror of deepmr's Sampling, _sparse/plan.py:97-126).

    ``index`` holds the per-axis (nframes, npts) integer indexes in (z, y, x)
    order (host tensors); ``plan`` is the width-1 device plan that computes.
    """

    index: tuple
    dshape: tuple
    ishape: tuple
    ndim: int
    device: str
    plan: object = field(default=None, repr=False)

    def to(self, device):
        if torch.device(device).type != "cpu":
            self.plan.to(device)
            self.device = self.plan.device
        return self

    @property
    def nframes(self):
        return int(self.index[0].shape[0])


def prepare_sampling(indexes, shape, device="cpu"):
    """Precompute a sparse sampling mask (src/deepmr/fft/sparse_fft.py:21-66,
    _sparse/plan.py:14-94). ``indexes`` (ncontrasts, nviews, nsamples, ndim),
    x first, integers in [0, shape); ``shape`` scalar or (z, y, x)."""
    idx = torch.as_tensor(indexes).detach().cpu().to(torch.int16).to(torch.int64)
    ndim = idx.shape[-1]
    if ndim not in (1, 2, 3):
        raise ValueError("sparse FFT masks are 1D, 2D or 3D")
    shape = [int(shape)] * ndim if np.isscalar(shape) else [int(s) for s in list(shape)[-ndim:]]
    ishape = list(idx.shape[:-1])
    while len(ishape) < 3:
        ishape = [1] + ishape
    nframes = int(ishape[0])
    ishape = tuple(int(s) for s in ishape[1:])
    npts = int(np.prod(ishape))
    idx = idx.reshape(nframes, npts, ndim)
    for d in range(ndim):  # d = 0 is x, the last grid axis
        n = shape[ndim - 1 - d]
        if npts and (int(idx[..., d].min()) < 0 or int(idx[..., d].max()) >= n):
            raise ValueError(f"sampling indexes along axis {d} (x first) outside [0, {n})")
    index = tuple(idx[..., d].contiguous() for d in range(ndim))[::-1]
    # device plan: coordinates are the indexes themselves (grid units)
    kdim = max(ndim, 2)
    coord = np.zeros((nframes, npts, kdim), np.float32)
    coord[..., :ndim] = idx.numpy().astype(np.float32)
    kshape = tuple(shape) if ndim > 1 else (1, shape[0])
    dev = _core.compute_device(device)
    plan = _nufft.NUFFTPlan(
        ndim=kdim, oversamp=(1.0,) * kdim, width=(1,) * kdim, beta=(0.0,) * kdim,
        os_shape=kshape, shape=kshape, interpolator=None, device=str(dev), ishape=ishape,
        nframes=nframes, scale=1.0, coord=coord,
    )
    plan.interpolator = _nufft._build_native(plan, dev, shift=True)
    return Sampling(index, tuple(shape), ishape, ndim, str(dev), plan)


def _unit_axis(x, ndim):
    """1D data rides on the 2D kernels with a unit y axis."""
    if ndim != 1:
        return x
    return x[..., None, :]


def _apply_sparse_fft(image, sampling_mask, basis_adjoint, weight, device, threadsperblock):
    mask = sampling_mask
    if device is not None:
        mask.to(device)
    out = _nufft._apply_nufft(_unit_axis(image, mask.ndim), mask.plan, basis_adjoint, weight, None,
                              threadsperblock)
    return out


def _apply_sparse_ifft(kspace, sampling_mask, basis, weight, device, threadsperblock):
    mask = sampling_mask
    if device is not None:
        mask.to(device)
    img = _nufft._apply_nufft_adj(kspace, mask.plan, basis, weight, None, threadsperblock)
    if mask.ndim == 1:
        img = img[..., 0, :]
    return img


def _nvox(mask):
    return float(np.prod(mask.dshape))


class ApplySparseFFT(torch.autograd.Function):
    """Sparse FFT with the sparse IFFT as its backward (sparse_fft.py:111-137).

    The autograd convention follows tilenufft.fft.set_autograd_mode: "exact"
    scales the backward by prod(shape) (the true VJP of the unnormalised
    forward); "deepmr" keeps deepmr's unscaled wiring.
    """

    @staticmethod
    def forward(image, sampling_mask, basis_adjoint, weight, device, threadsperblock):
        return _apply_sparse_fft(image, sampling_mask, basis_adjoint, weight, device, threadsperblock)

    @staticmethod
    def setup_context(ctx, inputs, output):
        _, sampling_mask, basis_adjoint, weight, device, threadsperblock = inputs
        ctx.set_materialize_grads(False)
        ctx.mask = sampling_mask
        ctx.basis_adjoint = basis_adjoint
        ctx.weight = weight
        ctx.device = device
        ctx.threadsperblock = threadsperblock

    @staticmethod
    def backward(ctx, kspace):
        if kspace is None:
            return (None,) * 6
        ba = ctx.basis_adjoint
        basis = None if ba is None else torch.as_tensor(ba).conj().t()
        grad = _apply_sparse_ifft(kspace, ctx.mask, basis, ctx.weight, ctx.device, ctx.threadsperblock)
        if _nufft.get_autograd_mode() == "exact":
            grad = grad * _nvox(ctx.mask)
        return grad, None, None, None, None, None


def apply_sparse_fft(image, sampling_mask, basis_adjoint=None, weight=None, device=None,
                     threadsperblock=128):
    """Sparse FFT (src/deepmr/fft/sparse_fft.py:140-179): image
    (..., [ncontrasts], *shape) -> k-space (..., [ncontrasts], nviews, nsamples).
    With ``basis_adjoint`` (ncontrasts, ncoeff) the image holds ncoeff
    subspace coefficients in place of the contrasts."""
    return ApplySparseFFT.apply(image, sampling_mask, basis_adjoint, weight, device, threadsperblock)


class ApplySparseIFFT(torch.autograd.Function):
    """Sparse IFFT with the sparse FFT as its backward (sparse_fft.py:182-220)."""

    @staticmethod
    def forward(kspace, sampling_mask, basis, weight, device, threadsperblock):
        return _apply_sparse_ifft(kspace, sampling_mask, basis, weight, device, threadsperblock)

    @staticmethod
    def setup_context(ctx, inputs, output):
        _, sampling_mask, basis, weight, device, threadsperblock = inputs
        ctx.set_materialize_grads(False)
        ctx.mask = sampling_mask
        ctx.basis = basis
        ctx.weight = weight
        ctx.device = device
        ctx.threadsperblock = threadsperblock

    @staticmethod
    def backward(ctx, image):
        if image is None:
            return (None,) * 6
        b = ctx.basis
        basis_adjoint = None if b is None else torch.as_tensor(b).conj().t()
        grad = _apply_sparse_fft(image, ctx.mask, basis_adjoint, ctx.weight, ctx.device, ctx.threadsperblock)
        if _nufft.get_autograd_mode() == "exact":
            grad = grad / _nvox(ctx.mask)
        return grad, None, None, None, None, None


def apply_sparse_ifft(kspace, sampling_mask, basis=None, weight=None, device=None, threadsperblock=128):
    """Sparse IFFT (src/deepmr/fft/sparse_fft.py:223-258): k-space
    (..., [ncontrasts], nviews, nsamples) -> image; ``basis`` (ncoeff,
    ncontrasts) returns ncoeff subspace coefficients."""
    return ApplySparseIFFT.apply(kspace, sampling_mask, basis, weight, device, threadsperblock)


@dataclass
class SparseGramMatrix:
    """Toeplitz kernel of the sparse FFT (mirror of _sparse/toeplitz.py:214-238)."""

    value: torch.Tensor
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
        if self._colsum is None or self._colsum.device != self.value.device:
            self._colsum = self.value.sum(dim=1).contiguous()
        return self._colsum


def plan_toeplitz_fft(coord, shape, basis=None, device="cpu", lowrank_product="deepmr"):
    """Toeplitz kernel of the sparse FFT (src/deepmr/fft/sparse_fft.py:69-94,
    _sparse/toeplitz.py:18-114): the scattered sampling mask (with the
    subspace contraction for a ``basis`` (ncoeff, ncontrasts)), in FFT order,
    normalised by its mean non-zero magnitude. ``lowrank_product`` as in
    plan_toeplitz_nufft."""
    if lowrank_product not in ("deepmr", "matvec"):
        raise ValueError("lowrank_product must be 'deepmr' or 'matvec'")
    idx = torch.as_tensor(coord).detach().cpu().to(torch.int32)
    ndim = idx.shape[-1]
    isreal = False
    if basis is not None:
        basis = torch.as_tensor(basis).cpu()
        isreal = not torch.is_complex(basis)
        basis = basis.to(torch.float32 if isreal else torch.complex64).resolve_conj()
        ncoeff = basis.shape[0]
        adj = basis.conj().T
        delta = torch.ones((ncoeff, *idx.shape[:-1]), dtype=adj.dtype)
        rev = list(range(delta.dim() - 1, -1, -1))
        delta = (delta.permute(*rev) * adj).permute(*rev)
    else:
        delta = torch.ones(idx.shape[:-1], dtype=torch.float32)
    mask = prepare_sampling(idx, shape, device)
    dev = torch.device(mask.device)
    plan = mask.plan
    y = delta.to(dev).to(torch.complex64).resolve_conj()
    # the device plan writes the scatter in FFT order, i.e. already
    # ifftshift-ed (_sparse/toeplitz.py:94)
    grid, batch_shape, nimgax = _nufft._grid_step(y, plan, basis, None, dev)
    if plan.nframes > 1 or basis is not None:
        kern = grid.reshape(*batch_shape, nimgax, *plan.os_shape)
    else:
        kern = grid.reshape(*batch_shape, *plan.os_shape)
    if ndim == 1:
        kern = kern[..., 0, :]
    if isreal:
        kern = kern.real
    if basis is not None:
        kern = kern.reshape(*kern.shape[:2], -1).permute(2, 1, 0).contiguous()
    kern = kern / torch.mean(torch.abs(kern[kern != 0]))
    kern = torch.nan_to_num(kern)
    return SparseGramMatrix(kern, ndim, str(dev), basis is not None, lowrank_product)


def _selfadj(image, kern):
    isnumpy = isinstance(image, np.ndarray)
    x = torch.as_tensor(image)
    x = x.to(torch.float32) if x.dtype in (torch.float16, torch.float32, torch.float64) else x.to(torch.complex64)
    odevice = x.device
    x = x.to(torch.device(kern.device))
    ax = tuple(range(-kern.ndim, 0))
    k = torch.fft.fftn(x, dim=ax, norm="ortho")
    y = torch.fft.ifftn(_kspace_product(k, kern, kern.ndim), dim=ax, norm="ortho").to(odevice)
    return y.numpy(force=True) if isnumpy else y


class ApplySparseFFTSelfadjoint(torch.autograd.Function):
    """F^H M F through the mask kernel; self-adjoint, so its own backward."""

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


def apply_sparse_fft_selfadj(image, toeplitz_kern, device=None, threadsperblock=128):
    """Self-adjoint sparse FFT via the Toeplitz kernel (sparse_fft.py:261-292, :459-503)."""
    if device is not None:
        toeplitz_kern.to(device)
    return ApplySparseFFTSelfadjoint.apply(image, toeplitz_kern, device, threadsperblock)
