"""NUFFT plan / apply on MI355X (drop-in for deepmr.fft.nufft's module).

Mirrors ``src/deepmr/fft/nufft.py``: ``plan_nufft``, ``apply_nufft``,
``apply_nufft_adj``, ``NUFFTPlan`` and the autograd Functions, with the same
argument meaning, tensor layouts, dtype casts and numpy-in / numpy-out
behaviour. The compute chain per direction is three device steps:

  forward  tn_fft_forward (apodize, pad, pruned 2D DFT) -> tn_interp
  adjoint  tn_grid -> tn_fft_adjoint (pruned inverse 2D DFT, crop, apodize)
  (plans without the pruned FFT stages, e.g. 3D: tn_apod_pad / tn_crop_apod
  around torch.fft on the device)

The centring shifts of deepmr's ``fft``/``ifft`` and the zero-pad / crop are
folded into the index maps of the native kernels, and the ``1/scale`` of the
interpolator into the tap weights.
"""

__all__ = [
    "plan_nufft",
    "apply_nufft",
    "apply_nufft_adj",
    "NUFFTPlan",
    "ApplyNUFFT",
    "ApplyNUFFTAdjoint",
    "set_autograd_mode",
    "get_autograd_mode",
    "retune",
]

import math

from dataclasses import dataclass, field

import numpy as np
import torch

from .. import _core

# --------------------------------------------------------------------------
# plan
# --------------------------------------------------------------------------


@dataclass
class NUFFTPlan:
    """Precomputed NUFFT (mirror of deepmr's NUFFTPlan, nufft.py:379-404).

    Attributes keep deepmr's names and (z, y, x) tuple order. ``interpolator``
    is the device plan (sorted trajectory and tile bins) instead of tap tables.
    """

    ndim: int
    oversamp: tuple
    width: tuple
    beta: tuple
    os_shape: tuple
    shape: tuple
    interpolator: object
    device: str
    ishape: tuple = ()
    nframes: int = 1
    scale: float = 1.0
    coord: np.ndarray = field(default=None, repr=False)  # scaled, (nframes, npts, ndim)
    tile: tuple = None
    interp_chunk: int = 0
    grid_chunk: int = 0
    itile: tuple = None
    interp_mode: int = 0
    grid_split: int = 0
    grid_mode: int = 0
    lowrank_native: object = field(default=None, repr=False, compare=False)

    def to(self, device):
        """Move the device plan to another HIP device ("cpu" keeps it where it is)."""
        target = _core.compute_device(device) if torch.device(device).type != "cpu" else None
        if target is not None and target != self.interpolator.device:
            self.interpolator = _build_native(self, target, shift=True)
            self.lowrank_native = None
            self.device = str(target)
        return self

    @property
    def npts(self):
        return int(np.prod(self.ishape))


def _build_native(plan, device, shift, coord=None):
    ndim = plan.ndim
    sh = [(g - g // 2) % g for g in plan.os_shape] if shift else [0] * ndim  # unit axes: 0
    apod = [
        _core.apodization(plan.shape[d], plan.oversamp[d], plan.width[d], plan.beta[d])
        if plan.width[d] != 1
        else None
        for d in range(ndim)
    ]
    return _core.NativePlan(
        plan.coord if coord is None else coord, plan.shape, plan.os_shape, plan.width, plan.beta,
        1.0 / plan.scale, sh,
        device, apod=apod, tile=plan.tile, interp_chunk=plan.interp_chunk,
        grid_chunk=plan.grid_chunk, itile=plan.itile, interp_mode=plan.interp_mode,
        grid_split=plan.grid_split, grid_mode=plan.grid_mode,
    )


def retune(plan, tile=None, interp_chunk=None, grid_chunk=None, itile=None, interp_mode=None,
           grid_split=None, grid_mode=None):
    """Rebuild the device plan of ``plan`` with other launch parameters."""
    if grid_mode is not None:
        plan.grid_mode = int(grid_mode)
    if grid_split is not None:
        plan.grid_split = int(grid_split)
    if tile is not None:
        plan.tile = tuple(int(t) for t in tile)
    if itile is not None:
        plan.itile = tuple(int(t) for t in itile)
    if interp_mode is not None:
        plan.interp_mode = int(interp_mode)
    if interp_chunk is not None:
        plan.interp_chunk = int(interp_chunk)
    if grid_chunk is not None:
        plan.grid_chunk = int(grid_chunk)
    plan.interpolator = _build_native(plan, plan.interpolator.device, shift=True)
    plan.lowrank_native = None
    return plan


def _lowrank_native(plan):
    """Device plan holding every frame's samples as one trajectory (built on
    first low-rank use): the subspace kernels read frame f's samples at
    [f * npts, (f + 1) * npts) and its basis row from the sample index."""
    dev = plan.interpolator.device
    nat = plan.lowrank_native
    if nat is None or nat.device != dev:
        merged = np.ascontiguousarray(plan.coord.reshape(1, -1, plan.ndim))
        nat = _build_native(plan, dev, shift=True, coord=merged)
        plan.lowrank_native = nat
    return nat


def _basis_tensor(basis, rows, cols, name, dev):
    """deepmr casts the basis to float32 or complex64 (nufft.py:466-474,
    :547-555); the kernels take it as complex64 (rows, cols) on the device."""
    b = torch.as_tensor(basis)
    if b.dim() != 2 or tuple(b.shape) != (rows, cols):
        raise ValueError(f"{name} must be ({rows}, {cols}), got {tuple(b.shape)}")
    if b.dtype in (torch.float16, torch.float32, torch.float64, torch.bfloat16):
        b = b.to(torch.float32)
    return b.to(dev).to(torch.complex64).resolve_conj().contiguous()


def _expand_coord(coord):
    """(..., ndim) -> (nframes, npts, ndim) plus deepmr's ishape (plan.py:66-77)."""
    ishape = list(coord.shape[:-1])
    while len(ishape) < 3:
        ishape = [1] + ishape
    nframes = int(ishape[0])
    ishape = tuple(int(s) for s in ishape[1:])
    return coord.reshape(nframes, int(np.prod(ishape)), coord.shape[-1]), nframes, ishape


def plan_nufft(coord, shape, width=4, oversamp=1.25, device="cpu", tile=None, interp_chunk=0,
               grid_chunk=0, itile=None, interp_mode=0, grid_split=0, grid_mode=0):
    """Precompute a NUFFT (deepmr ``plan_nufft``, src/deepmr/fft/nufft.py:27-132).

    Parameters follow deepmr: ``coord`` (ncontrasts, nviews, nsamples, ndim)
    in [-N/2, N/2) with x first; ``shape``, ``width``, ``oversamp`` scalar or
    (z, y, x). ``device`` selects the HIP device ("cpu" = current HIP device).
    ``tile`` / ``itile`` (gridding / staged-interpolation tiles, (z, y, x)),
    ``interp_mode`` (1 gather, 2 LDS-staged), ``interp_chunk`` and
    ``grid_chunk`` (coils per work item), ``grid_split`` (max samples per
    gridding workgroup before a hot tile is split; < 0 never) and
    ``grid_mode`` (1 per-sample LDS atomics, 2 register runs over
    origin-sorted samples, 3 per-cell sums over a plan-time sparse gridding
    matrix, the automatic choice when that matrix fits in 4 GiB, 4 (3D,
    widths <= 4, grids of whole 8x8x4 tiles) one wave per tile summing its
    cells in registers, the automatic choice when the matrix does not fit)
    optionally override the launch parameters (0 / None = automatic).
    """
    coord = torch.as_tensor(coord).detach().cpu().to(torch.float32).numpy().copy()
    ndim = coord.shape[-1]
    width = np.asarray([width] * ndim if np.isscalar(width) else list(width), dtype=np.int16)
    oversamp = np.asarray([oversamp] * ndim if np.isscalar(oversamp) else list(oversamp),
                          dtype=np.float32)
    beta = _core.kaiser_bessel_beta(width, oversamp)
    if np.isscalar(shape):
        shape = np.asarray([shape] * ndim, dtype=np.int16)
    else:
        shape = np.asarray(list(shape), dtype=np.int16)[-ndim:]
    is_cart = _core.cartesian_axes(coord, shape)
    oversamp[is_cart] = 1.0
    width[is_cart] = 1
    os_shape = _core.oversampled_shape(shape, oversamp)
    scaled = _core.scale_coordinates(coord, shape, oversamp)
    cart_x = is_cart[::-1]
    for d in range(ndim):  # deepmr reads Cartesian taps at int(g) without wrap
        if cart_x[d]:
            gi = scaled[..., d].astype(np.int64)
            if gi.size and (gi.min() < 0 or gi.max() >= os_shape[ndim - 1 - d]):
                raise ValueError("Cartesian coordinates outside the grid")
    flat, nframes, ishape = _expand_coord(scaled)
    scale = _core.kernel_scale(width, beta)
    dev = _core.compute_device(device)
    plan = NUFFTPlan(
        ndim=ndim,
        oversamp=tuple(oversamp),
        width=tuple(int(w) for w in width),
        beta=tuple(beta),
        os_shape=tuple(int(s) for s in os_shape),
        shape=tuple(int(s) for s in shape),
        interpolator=None,
        device=str(dev),
        ishape=ishape,
        nframes=nframes,
        scale=scale,
        coord=flat,
        tile=None if tile is None else tuple(int(t) for t in tile),
        interp_chunk=int(interp_chunk),
        grid_chunk=int(grid_chunk),
        itile=None if itile is None else tuple(int(t) for t in itile),
        interp_mode=int(interp_mode),
        grid_split=int(grid_split),
        grid_mode=int(grid_mode),
    )
    plan.interpolator = _build_native(plan, dev, shift=True)
    return plan


# --------------------------------------------------------------------------
# autograd
# --------------------------------------------------------------------------

_AUTOGRAD_MODE = "exact"


def set_autograd_mode(mode):
    """Select the backward convention of ApplyNUFFT / ApplyNUFFTAdjoint.

    "exact"  : backward of the forward A is A^H and backward of deepmr's
               adjoint B = A^H / prod(os_shape) is B^H (true VJPs).
    "deepmr" : deepmr's wiring (src/deepmr/fft/nufft.py:189-209, :266-286):
               each backward calls the other apply function, which makes the
               gradients prod(os_shape) (or 1/prod) times the true VJP.
    Gradients through A^H A compositions (CG-SENSE) agree in both modes.
    """
    global _AUTOGRAD_MODE
    if mode not in ("exact", "deepmr"):
        raise ValueError("mode must be 'exact' or 'deepmr'")
    _AUTOGRAD_MODE = mode


def get_autograd_mode():
    return _AUTOGRAD_MODE


def _nos(plan):
    return float(np.prod(plan.os_shape))


class ApplyNUFFT(torch.autograd.Function):
    """Forward NUFFT with the hand-written adjoint as its backward."""

    @staticmethod
    def forward(image, nufft_plan, basis_adjoint, weight, device, threadsperblock):
        return _apply_nufft(image, nufft_plan, basis_adjoint, weight, device, threadsperblock)

    @staticmethod
    def setup_context(ctx, inputs, output):
        _, nufft_plan, basis_adjoint, weight, device, threadsperblock = inputs
        ctx.set_materialize_grads(False)
        ctx.nufft_plan = nufft_plan
        ctx.basis_adjoint = basis_adjoint
        ctx.weight = weight
        ctx.device = device
        ctx.threadsperblock = threadsperblock

    @staticmethod
    def backward(ctx, kspace):
        if kspace is None:
            return (None,) * 6
        basis = None if ctx.basis_adjoint is None else ctx.basis_adjoint.conj().t()
        grad = _apply_nufft_adj(kspace, ctx.nufft_plan, basis, ctx.weight, ctx.device,
                                ctx.threadsperblock)
        if _AUTOGRAD_MODE == "exact":
            grad = grad * _nos(ctx.nufft_plan)
        return grad, None, None, None, None, None


class ApplyNUFFTAdjoint(torch.autograd.Function):
    """deepmr's adjoint NUFFT with the hand-written forward as its backward."""

    @staticmethod
    def forward(kspace, nufft_plan, basis, weight, device, threadsperblock):
        return _apply_nufft_adj(kspace, nufft_plan, basis, weight, device, threadsperblock)

    @staticmethod
    def setup_context(ctx, inputs, output):
        _, nufft_plan, basis, weight, device, threadsperblock = inputs
        ctx.set_materialize_grads(False)
        ctx.nufft_plan = nufft_plan
        ctx.basis = basis
        ctx.weight = weight
        ctx.device = device
        ctx.threadsperblock = threadsperblock

    @staticmethod
    def backward(ctx, image):
        if image is None:
            return (None,) * 6
        basis_adjoint = None if ctx.basis is None else ctx.basis.conj().t()
        grad = _apply_nufft(image, ctx.nufft_plan, basis_adjoint, ctx.weight, ctx.device,
                            ctx.threadsperblock)
        if _AUTOGRAD_MODE == "exact":
            grad = grad / _nos(ctx.nufft_plan)
        return grad, None, None, None, None, None


def apply_nufft(image, nufft_plan, basis_adjoint=None, weight=None, device=None, threadsperblock=128):
    """Apply the NUFFT (deepmr ``apply_nufft``, src/deepmr/fft/nufft.py:212-245).

    image: (..., ncontrasts, ny, nx) or (..., ncontrasts, nz, ny, nx);
    returns k-space (..., ncontrasts, nviews, nsamples). ``threadsperblock``
    is accepted for compatibility; launch shapes come from the plan.
    """
    return ApplyNUFFT.apply(image, nufft_plan, basis_adjoint, weight, device, threadsperblock)


def apply_nufft_adj(kspace, nufft_plan, basis=None, weight=None, device=None, threadsperblock=128):
    """Apply deepmr's adjoint NUFFT (src/deepmr/fft/nufft.py:289-322).

    Like deepmr, this is the adjoint divided by prod(os_shape) (both FFTs use
    norm=None).
    """
    return ApplyNUFFTAdjoint.apply(kspace, nufft_plan, basis, weight, device, threadsperblock)


# --------------------------------------------------------------------------
# implementation
# --------------------------------------------------------------------------


def _as_compute(x):
    """deepmr's dtype rule: real -> float32 (computed as complex), else complex64."""
    x = torch.as_tensor(x)
    if x.dtype in (torch.float16, torch.float32, torch.float64, torch.bfloat16):
        return x.to(torch.float32)
    return x.to(torch.complex64).resolve_conj()  # lazy conj views carry unconjugated memory


def _fused_weight(weight, plan, dev, nframes_axis_shape):
    """Return (device weight tensor of (nframes, npts), None) when it can be
    fused into the kernels, else (None, broadcast weight)."""
    if weight is None:
        return None, None
    w = torch.as_tensor(weight, dtype=torch.float32)
    target = plan.nframes * plan.npts
    trailing = tuple(nframes_axis_shape)
    if w.numel() == target and tuple(w.shape) == trailing[len(trailing) - w.dim():]:
        return w.to(dev).reshape(plan.nframes, plan.npts).contiguous(), None
    return None, w.to(dev)


def _check_plan(nufft_plan, device):
    if getattr(nufft_plan, "interpolator", None) is None:
        raise ValueError("nufft_plan has no device interpolator")
    if device is not None:
        nufft_plan.to(device)
    return nufft_plan.interpolator.device


def _apply_nufft(image, nufft_plan, basis_adjoint, weight, device, threadsperblock):
    isnumpy = isinstance(image, np.ndarray)
    image = _as_compute(image)
    dev = _check_plan(nufft_plan, device)
    odevice = image.device
    plan = nufft_plan
    ndim, nframes = plan.ndim, plan.nframes
    x = image.to(dev)
    if x.dtype != torch.complex64:
        x = x.to(torch.complex64)
    lead = tuple(x.shape[:-ndim])
    if tuple(x.shape[-ndim:]) != tuple(plan.shape):
        raise ValueError(f"image spatial shape {tuple(x.shape[-ndim:])} != plan shape {plan.shape}")
    # low rank: the image holds ncoeff subspace coefficients where the plain
    # transform holds the nframes contrasts (interp.py:84-92)
    nimgax = nframes
    ba = None
    if basis_adjoint is not None:
        ncoeff = int(torch.as_tensor(basis_adjoint).shape[-1])
        ba = _basis_tensor(basis_adjoint, nframes, ncoeff, "basis_adjoint", dev)
        nimgax = ncoeff
        if nframes == 1 and ncoeff != 1:
            raise ValueError("a single-frame plan takes a (1, 1) basis_adjoint")
    if nframes == 1:
        batch_shape = lead
    else:
        if len(lead) < 1 or lead[-1] != nimgax:
            what = "subspace coefficients" if ba is not None else "contrasts"
            raise ValueError(f"expected {nimgax} {what} before the spatial axes, got {lead}")
        batch_shape = lead[:-1]
    nbatch = int(np.prod(batch_shape))
    nimg = nbatch * nimgax
    x = x.contiguous()
    native = plan.interpolator
    with torch.cuda.device(dev):
        grid = torch.empty((nimg, *plan.os_shape), dtype=torch.complex64, device=dev)
        if native.has_fft:
            k = native.fft_forward(x, grid, nimg)
        else:
            native.apod_pad(x, grid, nimg)
            k = torch.fft.fftn(grid, dim=tuple(range(-ndim, 0)))
        del grid
        ksp_shape = (*batch_shape, nframes, *plan.ishape) if nframes > 1 else (*batch_shape, *plan.ishape)
        w_fused, w_bcast = _fused_weight(weight, plan, dev, (nframes, *plan.ishape) if nframes > 1 else plan.ishape)
        out = torch.empty((nbatch, nframes, plan.npts), dtype=torch.complex64, device=dev)
        if nbatch > 0 and ba is not None:
            _lowrank_native(plan).interp_lowrank(k, out, nbatch, nimgax, plan.npts, ba, w_fused)
        elif nbatch > 0:
            native.interp(k, out, nbatch, w_fused)
        out = out.reshape(ksp_shape)
        if w_bcast is not None:
            out = w_bcast * out
    out = out.to(odevice)
    if isnumpy:
        out = out.numpy(force=True)
    return out


def _grid_step(y, plan, basis, weight, dev):
    """Weighting and gridding of k-space ``y`` onto the plan's (FFT-ordered)
    grid: returns (grid (nbatch, nimgax, *os_shape), batch_shape, nimgax).
    ``nimgax`` is nframes, or ncoeff for a subspace ``basis``; a single-frame
    plan keeps coefficient 0 (grid.py:117)."""
    ndim, nframes = plan.ndim, plan.nframes
    w_fused, w_bcast = _fused_weight(weight, plan, dev, (nframes, *plan.ishape) if nframes > 1 else plan.ishape)
    if w_bcast is not None:
        y = w_bcast * y
    batch_shape = tuple(y.shape[:-2]) if nframes == 1 else tuple(y.shape[:-3])
    nbatch = int(np.prod(batch_shape))
    if y.numel() != nbatch * nframes * plan.npts:
        raise ValueError(f"k-space shape {tuple(y.shape)} does not match the plan "
                         f"({nframes} frames x {plan.npts} samples)")
    y = y.reshape(nbatch, nframes, plan.npts).contiguous()
    # low rank: the image gets ncoeff subspace coefficients per batch entry
    # (grid.py:89-104)
    nimgax = nframes
    bs = None
    if basis is not None:
        ncoeff = int(torch.as_tensor(basis).shape[0])
        bs = _basis_tensor(basis, ncoeff, nframes, "basis", dev)
        nimgax = ncoeff
    with torch.cuda.device(dev):
        grid = torch.empty((nbatch, nimgax, *plan.os_shape), dtype=torch.complex64, device=dev)
        if nbatch > 0 and bs is not None:
            _lowrank_native(plan).gridding_lowrank(y, grid, nbatch, nimgax, plan.npts, bs, w_fused)
        elif nbatch > 0:
            plan.interpolator.gridding(y, grid, nbatch, w_fused)
        if bs is not None and nframes == 1:
            grid = grid[:, :1].contiguous()
            nimgax = 1
    return grid, batch_shape, nimgax


def _apply_nufft_adj(kspace, nufft_plan, basis, weight, device, threadsperblock):
    isnumpy = isinstance(kspace, np.ndarray)
    kspace = _as_compute(kspace)
    dev = _check_plan(nufft_plan, device)
    odevice = kspace.device
    plan = nufft_plan
    ndim, nframes = plan.ndim, plan.nframes
    y = kspace.to(dev)
    if y.dtype != torch.complex64:
        y = y.to(torch.complex64)
    grid, batch_shape, nimgax = _grid_step(y, plan, basis, weight, dev)
    nbatch = int(np.prod(batch_shape))
    with torch.cuda.device(dev):
        nimg = nbatch * nimgax
        img = torch.empty((nimg, *plan.shape), dtype=torch.complex64, device=dev)
        if plan.interpolator.has_fft:
            plan.interpolator.fft_adjoint(grid, img, nimg)
        else:
            img_g = torch.fft.ifftn(grid, dim=tuple(range(-ndim, 0)))
            plan.interpolator.crop_apod(img_g, img, nimg)
            del img_g
        del grid
        shape = (*batch_shape, nimgax, *plan.shape) if nframes > 1 else (*batch_shape, *plan.shape)
        img = img.reshape(shape)
    img = img.to(odevice)
    if isnumpy:
        img = img.numpy(force=True)
    return img
