"""Interpolator plan / apply (drop-in for ``deepmr.fft._interp``).

Mirrors src/deepmr/fft/_interp/{__init__,plan,interp,grid}.py: the same
signatures, default arguments and layout rules, on the native kernels with
the identity index map (no folded FFT shifts).
"""

__all__ = [
    "Interpolator",
    "plan_interpolator",
    "apply_interpolation",
    "apply_gridding",
    "interpolate",
    "gridding",
]

from dataclasses import dataclass, field

import numpy as np
import torch

from .. import _core


@dataclass
class Interpolator:
    """Device interpolator (mirror of deepmr's Interpolator, plan.py:147-187).

    ``native`` holds the sorted trajectory and tile bins; taps are evaluated
    in registers at apply time instead of being stored.
    """

    native: object
    dshape: tuple
    ishape: tuple
    scale: float
    ndim: int
    device: str
    nframes: int = 1
    coord: np.ndarray = field(default=None, repr=False)
    width: tuple = ()
    beta: tuple = ()
    lowrank_native: object = field(default=None, repr=False, compare=False)

    def to(self, device):
        if torch.device(device).type == "cpu":
            return self
        target = _core.compute_device(device)
        if target != self.native.device:
            self.native = _native(self.coord, self.dshape, self.width, self.beta, self.scale, target)
            self.lowrank_native = None
            self.device = str(target)
        return self

    def merged(self):
        """Device plan of every frame's samples as one trajectory (low-rank
        kernels; built on first use)."""
        nat = self.lowrank_native
        if nat is None or nat.device != self.native.device:
            flat = np.ascontiguousarray(self.coord.reshape(1, -1, self.ndim))
            nat = _native(flat, self.dshape, self.width, self.beta, self.scale, self.native.device)
            self.lowrank_native = nat
        return nat


def _native(coord, dshape, width, beta, scale, device):
    return _core.NativePlan(coord, dshape, dshape, width, beta, 1.0 / scale, [0] * len(dshape), device)


def plan_interpolator(coord, shape, width=2, beta=1.0, device="cpu"):
    """Plan (deepmr ``plan_interpolator``, src/deepmr/fft/_interp/plan.py:14-143).

    coord: (ncontrasts, nviews, nsamples, ndim) grid-unit positions, x first.
    shape / width / beta: scalar or (z, y, x).
    """
    coord = torch.as_tensor(coord, dtype=torch.float32).detach().cpu().numpy()
    ndim = coord.shape[-1]
    ishape = list(coord.shape[:-1])
    while len(ishape) < 3:
        ishape = [1] + ishape
    nframes = int(ishape[0])
    ishape = tuple(int(s) for s in ishape[1:])
    shape = [int(shape)] * ndim if np.isscalar(shape) else [int(s) for s in list(shape)[-ndim:]]
    width = [int(width)] * ndim if np.isscalar(width) else [int(w) for w in width]
    beta = [float(beta)] * ndim if np.isscalar(beta) else [float(b) for b in beta]
    beta = [float(np.float32(b)) for b in beta]
    scale = _core.kernel_scale(width, beta)
    flat = np.ascontiguousarray(coord.reshape(nframes, int(np.prod(ishape)), ndim))
    for d in range(ndim):  # width-1 axes index int(g) without wrap (plan.py:118-121)
        if width[ndim - 1 - d] == 1:
            gi = flat[..., d].astype(np.int64)
            if gi.size and (gi.min() < 0 or gi.max() >= shape[ndim - 1 - d]):
                raise ValueError("width-1 axis coordinates outside the grid")
    dev = _core.compute_device(device)
    native = _native(flat, tuple(shape), tuple(width), tuple(beta), scale, dev)
    return Interpolator(native, tuple(shape), ishape, scale, ndim, str(dev), nframes, flat,
                        tuple(width), tuple(beta))


def _as_compute(x):
    x = torch.as_tensor(x)
    if x.dtype in (torch.float16, torch.float32, torch.float64, torch.bfloat16):
        return x.to(torch.float32)
    return x.to(torch.complex64).resolve_conj()  # lazy conj views carry unconjugated memory


def _basis(b, rows, cols, name, dev):
    b = torch.as_tensor(b)
    if b.dim() != 2 or tuple(b.shape) != (rows, cols):
        raise ValueError(f"{name} must be ({rows}, {cols}), got {tuple(b.shape)}")
    return _as_compute(b).to(dev).to(torch.complex64).resolve_conj().contiguous()


def apply_interpolation(data_in, interpolator, basis_adjoint=None, device=None, threadsperblock=128):
    """Grid -> samples (src/deepmr/fft/_interp/interp.py:14-120), incl. /scale.

    data_in (..., ncontrasts, *dshape) -> (..., ncontrasts, nviews, nsamples).
    """
    data_in = _as_compute(data_in)
    if device is not None:
        interpolator.to(device)
    native = interpolator.native
    dev = native.device
    odevice = data_in.device
    ndim, nframes = interpolator.ndim, interpolator.nframes
    x = data_in.to(dev).to(torch.complex64)
    batch_shape = tuple(x.shape[:-ndim]) if nframes == 1 else tuple(x.shape[: -ndim - 1])
    nbatch = int(np.prod(batch_shape))
    npts = int(np.prod(interpolator.ishape))
    ba = None
    nimgax = nframes
    if basis_adjoint is not None:  # grid holds ncoeff coefficients (interp.py:268-357)
        ncoeff = int(torch.as_tensor(basis_adjoint).shape[-1])
        ba = _basis(basis_adjoint, nframes, ncoeff, "basis_adjoint", dev)
        nimgax = ncoeff
    x = x.reshape(nbatch, nimgax, *interpolator.dshape).contiguous()
    with torch.cuda.device(dev):
        out = torch.empty((nbatch, nframes, npts), dtype=torch.complex64, device=dev)
        if nbatch > 0 and ba is not None:
            interpolator.merged().interp_lowrank(x, out, nbatch, nimgax, npts, ba)
        elif nbatch > 0:
            native.interp(x, out, nbatch)
    if nframes == 1:
        out = out.reshape(*batch_shape, *interpolator.ishape)
    else:
        out = out.reshape(*batch_shape, nframes, *interpolator.ishape)
    return out.to(odevice)


def apply_gridding(data_in, interpolator, basis=None, device=None, threadsperblock=128):
    """Samples -> grid (src/deepmr/fft/_interp/grid.py:14-125), incl. /scale.

    data_in (..., ncontrasts, nviews, nsamples) -> (..., ncontrasts, *dshape).
    """
    data_in = _as_compute(data_in)
    if device is not None:
        interpolator.to(device)
    native = interpolator.native
    dev = native.device
    odevice = data_in.device
    nframes = interpolator.nframes
    y = data_in.to(dev).to(torch.complex64)
    batch_shape = tuple(y.shape[:-2]) if nframes == 1 else tuple(y.shape[:-3])
    nbatch = int(np.prod(batch_shape))
    npts = int(np.prod(interpolator.ishape))
    y = y.reshape(nbatch, nframes, npts).contiguous()
    bs = None
    nimgax = nframes
    if basis is not None:  # ncoeff coefficient grids (grid.py:326-417)
        ncoeff = int(torch.as_tensor(basis).shape[0])
        bs = _basis(basis, ncoeff, nframes, "basis", dev)
        nimgax = ncoeff
    with torch.cuda.device(dev):
        out = torch.empty((nbatch, nimgax, *interpolator.dshape), dtype=torch.complex64, device=dev)
        if nbatch > 0 and bs is not None:
            interpolator.merged().gridding_lowrank(y, out, nbatch, nimgax, npts, bs)
        elif nbatch > 0:
            native.gridding(y, out, nbatch)
    if nframes == 1:  # a single-frame plan keeps coefficient 0 (grid.py:117)
        out = out[:, 0].reshape(*batch_shape, *interpolator.dshape)
    else:
        out = out.reshape(*batch_shape, nimgax, *interpolator.dshape)
    return out.to(odevice)


def interpolate(data_in, coord, shape=None, basis_adjoint=None, device="cpu", threadsperblock=128,
                width=2, beta=1.0):
    """Plan + interpolate (src/deepmr/fft/_interp/__init__.py:15-88)."""
    ndim = coord.shape[-1]
    if shape is None:
        shape = data_in.shape[-ndim:]
    interp = plan_interpolator(coord, shape, width, beta, device)
    return apply_interpolation(data_in, interp, basis_adjoint, threadsperblock=threadsperblock)


def gridding(data_in, coord, shape, basis=None, device="cpu", threadsperblock=128, width=2, beta=1.0):
    """Plan + grid (src/deepmr/fft/_interp/__init__.py:91-156)."""
    interp = plan_interpolator(coord, shape, width, beta, device)
    return apply_gridding(data_in, interp, basis, threadsperblock=threadsperblock)
