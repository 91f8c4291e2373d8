"""ORACLE — test infrastructure only.

CPU restatement of deepmr's Kaiser-Bessel NUFFT (``deepmr.fft.nufft`` /
``nufft_adj``) used as the checker for the HIP product path. Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
it; the product package never does.

Every function cites the reference file:line it restates (paths relative to
the deepmr repository, ``src/deepmr/`` abbreviated ``dm/``). The hot loops
(tap tables, interpolation, gridding) run in ``oracle/nufft_loops.c``
(built by ``oracle/Makefile`` into ``oracle/lib``); the glue (plan, layouts,
apodization, pad/crop, centred FFT) is numpy.

Pins (tests/test_oracle_*.py):
  * KB taps: A&S polynomial vs scipy.special.i0, brute-force tap windows.
  * scale: closed form and the on-grid invariant interp(ones) == 1.
  * resize: the reference's known answers (tests/signal/test_resize.py).
  * FFT conventions and the Cartesian width-1 path: the reference's
    known answers (tests/fft/test_nufft.py: centred delta <-> all ones).
  * full NUFFT: FP64 direct NDFT (oracle/ndft.py), dot-product tests.
"""

import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}
_FAST = False          # use the -ffast-math build of the loops (CPU baseline timing)
_FFT_BACKEND = "numpy"  # "torch": torch's CPU FFT, as deepmr's CPU path uses


def use_fast_loops(flag):
    """Route the raw loops through liborc_nufft_fast.so (baseline timing)."""
    global _FAST
    _FAST = bool(flag)


def set_fft_backend(name):
    """FFT used by fft_c / ifft_c: "numpy" (checker) or "torch" (deepmr's CPU FFT)."""
    global _FFT_BACKEND
    if name not in ("numpy", "torch"):
        raise ValueError(name)
    _FFT_BACKEND = name


def load_loops(fast=False):
    """Load the compiled C loops (``fast`` = the -ffast-math baseline build)."""
    name = "liborc_nufft_fast.so" if fast else "liborc_nufft.so"
    if name in _LIBS:
        return _LIBS[name]
    path = os.path.join(_HERE, "lib", name)
    if not os.path.exists(path):
        raise RuntimeError(
            f"oracle library {path} is missing; run `make -C oracle` first"
        )
    lib = ctypes.CDLL(path)
    f, i64, i32, P = ctypes.c_float, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p
    lib.orc_kaiser_bessel.restype = ctypes.c_double
    lib.orc_kaiser_bessel.argtypes = [ctypes.c_double, ctypes.c_double]
    lib.orc_prepare_interpolator.argtypes = [P, P, P, i64, i32, f, i32]
    lib.orc_interp2.argtypes = [P, P, P, P, P, P, i64, i64, i64, i64, i64, i32, i32]
    lib.orc_interp3.argtypes = [P] * 8 + [i64] * 6 + [i32] * 3
    lib.orc_grid2.argtypes = [P, P, P, P, P, P, i64, i64, i64, i64, i64, i32, i32]
    lib.orc_grid3.argtypes = [P] * 8 + [i64] * 6 + [i32] * 3
    lib.orc_interp_lowrank.argtypes = [P] * 9 + [i64] * 7 + [i32] * 3
    lib.orc_grid_lowrank.argtypes = [P] * 9 + [i64] * 7 + [i32] * 3
    lib.orc_set_threads.argtypes = [i32]
    lib.orc_get_threads.restype = i32
    _LIBS[name] = lib
    return lib


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


# --------------------------------------------------------------------------
# plan (dm/fft/nufft.py:27-132, dm/fft/_interp/plan.py:14-286)
# --------------------------------------------------------------------------
def kb_beta(width, oversamp):
    """Beatty's Kaiser-Bessel beta (dm/fft/nufft.py:98), float32 like deepmr."""
    width = np.asarray(width, dtype=np.int16)
    oversamp = np.asarray(oversamp, dtype=np.float32)
    return (
        np.float32(math.pi)
        * (((width / oversamp) * (oversamp - np.float32(0.5))) ** 2 - np.float32(0.8))
        ** np.float32(0.5)
    ).astype(np.float32)


def kaiser_bessel(u, beta):
    """A&S 9.8.1/9.8.2 KB kernel, float64 (dm/fft/_interp/plan.py:216-248)."""
    u = np.asarray(u, dtype=np.float64)
    out = np.zeros_like(u)
    inside = np.abs(u) <= 1
    x = float(beta) * np.sqrt(np.clip(1 - u**2, 0, None))
    t = x / 3.75
    small = inside & (x < 3.75)
    ts = t[small]
    out[small] = (
        1
        + 3.5156229 * ts**2
        + 3.0899424 * ts**4
        + 1.2067492 * ts**6
        + 0.2659732 * ts**8
        + 0.0360768 * ts**10
        + 0.0045813 * ts**12
    )
    big = inside & (x >= 3.75)
    xb, tb = x[big], t[big]
    out[big] = (
        xb**-0.5
        * np.exp(xb)
        * (
            0.39894228
            + 0.01328592 * tb**-1
            + 0.00225319 * tb**-2
            - 0.00157565 * tb**-3
            + 0.00916281 * tb**-4
            - 0.02057706 * tb**-5
            + 0.02635537 * tb**-6
            - 0.01647633 * tb**-7
            + 0.00392377 * tb**-8
        )
    )
    return out


def kernel_scaling(beta, width):
    """Product over axes of the summed on-node taps (dm/fft/_interp/plan.py:190-213)."""
    total = 1.0
    for b, w in zip(beta, width):
        w = int(w)
        if w == 1:
            continue
        start = math.ceil(-w / 2)
        total *= float(
            np.sum(kaiser_bessel((start + np.arange(w)) / (w / 2), float(b)))
        )
    return total


def os_shape_of(shape, oversamp):
    """Oversampled grid shape ceil(osf*N) (dm/fft/nufft.py:407-408)."""
    return np.ceil(np.asarray(oversamp, np.float32) * np.asarray(shape)).astype(np.int16)


def scale_coord(coord, shape_xyz, oversamp_xyz):
    """coord*G/N + G//2 per axis, float32 (dm/fft/nufft.py:411-420)."""
    out = np.array(coord, dtype=np.float32, copy=True)
    ndim = out.shape[-1]
    for i in range(-ndim, 0):
        g = np.ceil(np.float32(oversamp_xyz[i]) * shape_xyz[i])
        scale = g / shape_xyz[i]
        shift = g // 2
        out[..., i] = (out[..., i] * np.float32(scale)).astype(np.float32)
        out[..., i] = (out[..., i] + np.float32(shift)).astype(np.float32)
    return out


def cartesian_axes(coord, shape_zyx):
    """deepmr's Cartesian-axis test, including its axis-order quirk.

    dm/fft/nufft.py:105-109: shape is indexed (z, y, x) while coord[..., ax] is
    (x, y, z); the list is then reversed to (z, y, x).
    """
    ndim = coord.shape[-1]
    res = []
    for ax in range(ndim):
        a = (np.float32(shape_zyx[ax]) * coord[..., ax]).astype(np.float32)
        res.append(bool(np.allclose(a, np.round(a))))
    return np.asarray(res[::-1])


def plan_interpolator(coord, shape, width, beta):
    """Tap tables per axis (dm/fft/_interp/plan.py:14-143).

    coord: (..., ndim) grid-unit coordinates (x, y, z order).
    shape/width/beta: (z, y, x) order. Returns a dict mirroring Interpolator.
    """
    coord = np.asarray(coord, dtype=np.float32)
    ndim = coord.shape[-1]
    ishape = list(coord.shape[:-1])
    while len(ishape) < 3:
        ishape = [1] + ishape
    nframes = ishape[0]
    ishape = ishape[1:]
    npts = int(np.prod(ishape))
    shape = list(np.broadcast_to(shape, (ndim,)) if np.isscalar(shape) else shape[-ndim:])
    width = list(np.broadcast_to(width, (ndim,)))
    beta = list(np.broadcast_to(beta, (ndim,)))
    shape_x, width_x, beta_x = shape[::-1], width[::-1], beta[::-1]
    scale = kernel_scaling(beta_x, width_x)
    c = coord.reshape(nframes * npts, ndim).T
    lib = load_loops()
    index, value = [], []
    for i in range(ndim):  # (x, y, z)
        w = int(width_x[i])
        ci = np.ascontiguousarray(c[i], dtype=np.float32)
        if w == 1:  # plan.py:118-121 (truncation, no wrap)
            idx = ci.astype(np.int32)[:, None]
            val = np.ones((ci.size, 1), np.float32)
        else:
            val = np.zeros((ci.size, w), np.float32)
            idx = np.zeros((ci.size, w), np.int32)
            lib.orc_prepare_interpolator(
                _ptr(val), _ptr(idx), _ptr(ci), ci.size, w,
                np.float32(beta_x[i]), int(shape_x[i]),
            )
        index.append(idx.reshape(nframes, npts, w))
        value.append(val.reshape(nframes, npts, w))
    return dict(
        index=index[::-1], value=value[::-1], dshape=tuple(int(s) for s in shape),
        ishape=tuple(ishape), scale=scale, ndim=ndim,
    )


def plan_nufft(coord, shape, width=4, oversamp=1.25):
    """dm/fft/nufft.py:27-132."""
    coord = np.array(coord, dtype=np.float32, copy=True)
    ndim = coord.shape[-1]
    width = np.asarray([width] * ndim if np.isscalar(width) else width, np.int16)
    oversamp = np.asarray(
        [oversamp] * ndim if np.isscalar(oversamp) else oversamp, np.float32
    )
    beta = kb_beta(width, oversamp)
    shape = np.asarray([shape] * ndim if np.isscalar(shape) else list(shape)[-ndim:], np.int16)
    is_cart = cartesian_axes(coord, shape)
    oversamp[is_cart] = 1.0
    width[is_cart] = 1
    os_shape = os_shape_of(shape, oversamp)
    coord = scale_coord(coord, shape[::-1], oversamp[::-1])
    interp = plan_interpolator(coord, os_shape, width, beta)
    return dict(
        ndim=ndim, oversamp=tuple(oversamp), width=tuple(int(w) for w in width),
        beta=tuple(beta), os_shape=tuple(int(s) for s in os_shape),
        shape=tuple(int(s) for s in shape), interpolator=interp, scaled_coord=coord,
    )


# --------------------------------------------------------------------------
# pre/post processing
# --------------------------------------------------------------------------
def apodization_vector(n, oversamp, width, beta):
    """Per-axis apodization, float32 torch-like arithmetic (dm/fft/nufft.py:423-447)."""
    os_i = np.ceil(np.float32(oversamp) * n)
    idx = np.arange(n, dtype=np.float32)
    arg = (np.float32(beta) ** 2 - (np.float32(math.pi * int(width)) * (idx - n // 2) / np.float32(os_i)) ** 2).astype(np.float32)
    with np.errstate(invalid="ignore", divide="ignore"):
        apod = np.sqrt(arg.astype(np.float32))
        apod = (apod / np.sinh(apod)).astype(np.float32)
        apod = (apod / apod[int(n // 2)]).astype(np.float32)
    return np.nan_to_num(apod, nan=1.0).astype(np.float32)


def apodize(data, ndim, oversamp, width, beta):
    """In-place apodization along the last ndim axes (dm/fft/nufft.py:423-447)."""
    for n in range(1, ndim + 1):
        axis = -n
        if width[axis] != 1:
            i = data.shape[axis]
            apod = apodization_vector(i, oversamp[axis], width[axis], beta[axis])
            data *= apod.reshape([i] + [1] * (-axis - 1))
    return data


def resize(x, oshape):
    """Centred zero-pad / crop (dm/_signal/resize.py:11-102, :183-194)."""
    x = np.asarray(x)
    oshape = [oshape] if isinstance(oshape, int) else list(oshape)
    ishape = list(x.shape)
    nd = max(len(ishape), len(oshape))
    ie = [1] * (nd - len(ishape)) + ishape
    oe = [1] * (nd - len(oshape)) + oshape
    mx = np.maximum(ie, oe)
    ishape1 = list(mx[: nd - len(ishape)]) + ishape
    oshape1 = list(mx[: nd - len(oshape)]) + oshape
    if ishape1 == oshape1:
        return x
    ishift = [max(i // 2 - o // 2, 0) for i, o in zip(ishape1, oshape1)]
    oshift = [max(o // 2 - i // 2, 0) for i, o in zip(ishape1, oshape1)]
    copy = [min(i - si, o - so) for i, si, o, so in zip(ishape1, ishift, oshape1, oshift)]
    isl = tuple(slice(s, s + c) for s, c in zip(ishift, copy))
    osl = tuple(slice(s, s + c) for s, c in zip(oshift, copy))
    out = np.zeros(oshape1, dtype=x.dtype)
    out[osl] = x.reshape(ishape1)[isl]
    return out


def _torch_fft(x, ax, inverse):
    import torch

    t = torch.from_numpy(np.ascontiguousarray(x))
    t = torch.fft.ifftshift(t, dim=ax)
    t = torch.fft.ifftn(t, dim=ax) if inverse else torch.fft.fftn(t, dim=ax)
    return torch.fft.fftshift(t, dim=ax).numpy()


def fft_c(x, axes, norm=None):
    """Centred FFT, ifftshift -> fftn -> fftshift (dm/fft/fft.py:9-87)."""
    ax = tuple(a % x.ndim for a in sorted(axes))
    if _FFT_BACKEND == "torch" and norm is None:
        return _torch_fft(x, ax, False)
    y = np.fft.fftn(np.fft.ifftshift(x, axes=ax), axes=ax, norm=norm)
    return np.fft.fftshift(y, axes=ax)


def ifft_c(x, axes, norm=None):
    """Centred IFFT (dm/fft/fft.py:90-167); norm=None scales by 1/prod(G)."""
    ax = tuple(a % x.ndim for a in sorted(axes))
    if _FFT_BACKEND == "torch" and norm is None:
        return _torch_fft(x, ax, True)
    y = np.fft.ifftn(np.fft.ifftshift(x, axes=ax), axes=ax, norm=norm)
    return np.fft.fftshift(y, axes=ax)


# --------------------------------------------------------------------------
# interpolation / gridding with deepmr's layout rules
# --------------------------------------------------------------------------
def _tables(interp, ndim):
    idx = [np.ascontiguousarray(a, np.int32) for a in interp["index"]]
    val = [np.ascontiguousarray(a, np.float32) for a in interp["value"]]
    return idx, val


def _interp1(grid, idx, val):
    """dm/fft/_interp/interp.py:125-151 (_interpolate1), vectorised over the
    points: out[f, b, p] = sum_i val[f, p, i] * grid[f, b, idx[f, p, i]],
    taps added in order as the reference's loop does."""
    nframes, batch, _ = grid.shape
    out = np.zeros((nframes, batch, idx.shape[1]), np.complex64)
    for f in range(nframes):
        for i in range(idx.shape[-1]):
            out[f] += (val[f, :, i][None, :] * grid[f][:, idx[f, :, i]]).astype(np.complex64)
    return out


def _grid1(samp, idx, val, G):
    """dm/fft/_interp/grid.py:182-207 (_gridding1): cart[f, b, idx[f, p, i]] +=
    val[f, p, i] * samp[f, b, p] (np.add.at, tap-major rather than the
    reference's point-major order: the same sums up to float rounding)."""
    nframes, batch, _ = samp.shape
    out = np.zeros((nframes, batch, G), np.complex64)
    for f in range(nframes):
        for b in range(batch):
            for i in range(idx.shape[-1]):
                np.add.at(out[f, b], idx[f, :, i], (val[f, :, i] * samp[f, b]).astype(np.complex64))
    return out


def interpolate_raw(grid, interp, fast=None):
    """Kernel call on (nframes, batch, *G) -> (nframes, batch, npts).

    dm/fft/_interp/interp.py:125-227 (2D / 3D through oracle/nufft_loops.c,
    1D in numpy).
    """
    lib = load_loops(_FAST if fast is None else fast)
    ndim = interp["ndim"]
    idx, val = _tables(interp, ndim)
    grid = np.ascontiguousarray(grid, np.complex64)
    nframes, batch = grid.shape[:2]
    npts = idx[0].shape[1]
    G = grid.shape[2:]
    if nframes != idx[0].shape[0] or tuple(G) != tuple(interp["dshape"]):
        raise ValueError("grid shape does not match the interpolator")
    out = np.zeros((nframes, batch, npts), np.complex64)
    if ndim == 1:
        out = _interp1(grid, idx[0], val[0])
    elif ndim == 2:
        lib.orc_interp2(_ptr(out), _ptr(grid), _ptr(val[0]), _ptr(idx[0]), _ptr(val[1]),
                        _ptr(idx[1]), nframes, batch, npts, G[0], G[1],
                        idx[0].shape[-1], idx[1].shape[-1])
    elif ndim == 3:
        lib.orc_interp3(_ptr(out), _ptr(grid), _ptr(val[0]), _ptr(idx[0]), _ptr(val[1]),
                        _ptr(idx[1]), _ptr(val[2]), _ptr(idx[2]), nframes, batch, npts,
                        G[0], G[1], G[2], idx[0].shape[-1], idx[1].shape[-1],
                        idx[2].shape[-1])
    else:
        raise NotImplementedError("oracle covers 1D, 2D and 3D")
    return out


def grid_raw(samp, interp, fast=None):
    """Kernel call on (nframes, batch, npts) -> (nframes, batch, *G).

    dm/fft/_interp/grid.py:182-284 (2D / 3D through oracle/nufft_loops.c,
    1D in numpy).
    """
    lib = load_loops(_FAST if fast is None else fast)
    ndim = interp["ndim"]
    idx, val = _tables(interp, ndim)
    samp = np.ascontiguousarray(samp, np.complex64)
    nframes, batch, npts = samp.shape
    G = interp["dshape"]
    if nframes != idx[0].shape[0] or npts != idx[0].shape[1]:
        raise ValueError("sample shape does not match the interpolator")
    out = np.zeros((nframes, batch, *G), np.complex64)
    if ndim == 1:
        out = _grid1(samp, idx[0], val[0], G[0])
    elif ndim == 2:
        lib.orc_grid2(_ptr(out), _ptr(samp), _ptr(val[0]), _ptr(idx[0]), _ptr(val[1]),
                      _ptr(idx[1]), nframes, batch, npts, G[0], G[1],
                      idx[0].shape[-1], idx[1].shape[-1])
    elif ndim == 3:
        lib.orc_grid3(_ptr(out), _ptr(samp), _ptr(val[0]), _ptr(idx[0]), _ptr(val[1]),
                      _ptr(idx[1]), _ptr(val[2]), _ptr(idx[2]), nframes, batch, npts,
                      G[0], G[1], G[2], idx[0].shape[-1], idx[1].shape[-1],
                      idx[2].shape[-1])
    else:
        raise NotImplementedError("oracle covers 1D, 2D and 3D")
    return out


def _lowrank_args(interp):
    ndim = interp["ndim"]
    idx, val = _tables(interp, ndim)
    if ndim == 2:
        z_v, z_i = None, None
        (y_i, x_i), (y_v, x_v) = idx, val
        wz = 0
    else:
        (z_i, y_i, x_i), (z_v, y_v, x_v) = idx, val
        wz = z_i.shape[-1]
    return z_v, z_i, y_v, y_i, x_v, x_i, wz


def interpolate_lowrank_raw(grid, interp, basis_adjoint):
    """(ncoeff, batch, *G) -> (nframes, batch, npts) with basis_adjoint (nframes, ncoeff).

    dm/fft/_interp/interp.py:231-357 (1D: _interpolate_lowrank1, :231-265, in
    numpy: out[f, b, p] = sum_i val[f, p, i] sum_k ba[f, k] grid[k, b, idx]).
    """
    if interp["ndim"] == 1:
        idx, val = _tables(interp, 1)
        idx, val = idx[0], val[0]
        grid = np.asarray(grid, np.complex64)
        ba = np.asarray(basis_adjoint, np.complex64)
        nframes, npts = idx.shape[:2]
        out = np.zeros((nframes, grid.shape[1], npts), np.complex64)
        for f in range(nframes):
            mixed = np.tensordot(ba[f], grid, axes=(0, 0)).astype(np.complex64)  # (batch, G)
            for i in range(idx.shape[-1]):
                out[f] += (val[f, :, i][None, :] * mixed[:, idx[f, :, i]]).astype(np.complex64)
        return out
    lib = load_loops()
    z_v, z_i, y_v, y_i, x_v, x_i, wz = _lowrank_args(interp)
    grid = np.ascontiguousarray(grid, np.complex64)
    ba = np.ascontiguousarray(basis_adjoint, np.complex64)
    ncoeff, batch = grid.shape[:2]
    nframes, npts = y_i.shape[:2]
    G = list(grid.shape[2:])
    if len(G) == 2:
        G = [1] + G
    out = np.zeros((nframes, batch, npts), np.complex64)
    nul = ctypes.c_void_p(0)
    lib.orc_interp_lowrank(_ptr(out), _ptr(grid), _ptr(ba),
                           _ptr(z_v) if z_v is not None else nul,
                           _ptr(z_i) if z_i is not None else nul,
                           _ptr(y_v), _ptr(y_i), _ptr(x_v), _ptr(x_i),
                           nframes, ncoeff, batch, npts, G[0], G[1], G[2],
                           wz, y_i.shape[-1], x_i.shape[-1])
    return out


def grid_lowrank_raw(samp, interp, basis):
    """(nframes, batch, npts) -> (ncoeff, batch, *G) with basis (ncoeff, nframes).

    dm/fft/_interp/grid.py:288-417 (1D: _gridding_lowrank1, :288-322, in
    numpy: cart[k, b, idx] += val * basis[k, f] * samp[f, b, p]).
    """
    if interp["ndim"] == 1:
        idx, val = _tables(interp, 1)
        idx, val = idx[0], val[0]
        samp = np.asarray(samp, np.complex64)
        bs = np.asarray(basis, np.complex64)
        nframes, batch, _ = samp.shape
        G = interp["dshape"][0]
        out = np.zeros((bs.shape[0], batch, G), np.complex64)
        for k in range(bs.shape[0]):
            for f in range(nframes):
                g = _grid1((bs[k, f] * samp[f:f + 1]).astype(np.complex64), idx[f:f + 1], val[f:f + 1], G)
                out[k] += g[0]
        return out
    lib = load_loops()
    z_v, z_i, y_v, y_i, x_v, x_i, wz = _lowrank_args(interp)
    samp = np.ascontiguousarray(samp, np.complex64)
    bs = np.ascontiguousarray(basis, np.complex64)
    nframes, batch, npts = samp.shape
    ncoeff = bs.shape[0]
    G = list(interp["dshape"])
    Gf = [1] + G if len(G) == 2 else G
    out = np.zeros((ncoeff, batch, *G), np.complex64)
    nul = ctypes.c_void_p(0)
    lib.orc_grid_lowrank(_ptr(out), _ptr(samp), _ptr(bs),
                         _ptr(z_v) if z_v is not None else nul,
                         _ptr(z_i) if z_i is not None else nul,
                         _ptr(y_v), _ptr(y_i), _ptr(x_v), _ptr(x_i),
                         nframes, ncoeff, batch, npts, Gf[0], Gf[1], Gf[2],
                         wz, y_i.shape[-1], x_i.shape[-1])
    return out


def apply_interpolation(data_in, interp, basis_adjoint=None):
    """Layout rules of dm/fft/_interp/interp.py:14-120 (incl. the final /scale)."""
    data_in = np.asarray(data_in).astype(np.complex64)
    ndim, dshape, ishape = interp["ndim"], interp["dshape"], interp["ishape"]
    nframes = interp["index"][0].shape[0]
    npts = int(np.prod(ishape))
    batch_shape = data_in.shape[:-ndim] if nframes == 1 else data_in.shape[: -ndim - 1]
    batch = int(np.prod(batch_shape))
    if basis_adjoint is None:
        x = data_in.reshape(batch, nframes, *dshape).swapaxes(0, 1)
        out = interpolate_raw(x, interp)
    else:
        ncoeff = np.asarray(basis_adjoint).shape[-1]
        x = data_in.reshape(batch, ncoeff, *dshape).swapaxes(0, 1)
        out = interpolate_lowrank_raw(x, interp, basis_adjoint)
    if nframes == 1:
        out = out[0].reshape(*batch_shape, *ishape)
    else:
        out = out.swapaxes(0, 1).reshape(*batch_shape, nframes, *ishape)
    return (out / np.float32(interp["scale"])).astype(np.complex64)


def apply_gridding(data_in, interp, basis=None):
    """Layout rules of dm/fft/_interp/grid.py:14-125 (incl. the final /scale)."""
    data_in = np.asarray(data_in).astype(np.complex64)
    ndim, dshape, ishape = interp["ndim"], interp["dshape"], interp["ishape"]
    nframes = interp["index"][0].shape[0]
    npts = int(np.prod(ishape))
    batch_shape = data_in.shape[:-2] if nframes == 1 else data_in.shape[:-3]
    batch = int(np.prod(batch_shape))
    x = data_in.reshape(batch, nframes, npts).swapaxes(0, 1)
    if basis is None:
        ncoeff = nframes
        out = grid_raw(x, interp)
    else:
        ncoeff = np.asarray(basis).shape[0]
        out = grid_lowrank_raw(x, interp, basis)
    if nframes == 1:
        out = out[0].reshape(*batch_shape, *dshape)
    else:
        out = out.swapaxes(0, 1).reshape(*batch_shape, ncoeff, *dshape)
    return (out / np.float32(interp["scale"])).astype(np.complex64)


# --------------------------------------------------------------------------
# full transforms
# --------------------------------------------------------------------------
def apply_nufft(image, plan, weight=None, basis_adjoint=None):
    """Forward NUFFT chain (dm/fft/nufft.py:450-528)."""
    ndim = plan["ndim"]
    x = np.array(image, dtype=np.complex64, copy=True)
    apodize(x, ndim, plan["oversamp"], plan["width"], plan["beta"])
    x = resize(x, list(x.shape[:-ndim]) + list(plan["os_shape"]))
    k = fft_c(x, range(-ndim, 0)).astype(np.complex64)
    y = apply_interpolation(k, plan["interpolator"], basis_adjoint)
    if weight is not None:
        y = (np.asarray(weight, np.float32) * y).astype(np.complex64)
    return y


def apply_nufft_adj(kspace, plan, weight=None, basis=None):
    """Adjoint chain (dm/fft/nufft.py:531-606); note it is A^H / prod(os_shape)."""
    ndim = plan["ndim"]
    y = np.asarray(kspace).astype(np.complex64)
    if weight is not None:
        y = (np.asarray(weight, np.float32) * y).astype(np.complex64)
    g = apply_gridding(y, plan["interpolator"], basis)
    x = ifft_c(g, range(-ndim, 0)).astype(np.complex64)
    x = resize(x, list(x.shape[:-ndim]) + list(plan["shape"]))
    apodize(x, ndim, plan["oversamp"], plan["width"], plan["beta"])
    return x.astype(np.complex64)


def nufft(image, coord, shape=None, width=4, oversamp=1.25, basis_adjoint=None):
    """dm/fft/__init__.py:145-218."""
    ndim = np.asarray(coord).shape[-1]
    if shape is None:
        shape = np.asarray(image).shape[-ndim:]
    return apply_nufft(image, plan_nufft(coord, shape, width, oversamp),
                       basis_adjoint=basis_adjoint)


def nufft_adj(kspace, coord, shape, width=4, oversamp=1.25, basis=None):
    """dm/fft/__init__.py:221-286."""
    return apply_nufft_adj(kspace, plan_nufft(coord, shape, width, oversamp), basis=basis)


# --------------------------------------------------------------------------
# Toeplitz normal operator (dm/fft/toeplitz.py:15-157, dm/fft/nufft.py:609-669)
# --------------------------------------------------------------------------
def plan_toeplitz(coord, shape, dcf=None, width=4, basis=None):
    """PSF kernel on the 2x grid, dm/fft/toeplitz.py:15-157.

    With a subspace ``basis`` (ncoeff, nframes) the PSF of coefficient pair
    (c1 -> c2) is nufft_adj(dcf * conj(basis[c1, f]), basis)[c2] (delta at
    :97-101), and the kernel is returned as (voxels, c2, c1) (:143-147);
    a real basis keeps the real part only (:137-138).
    """
    coord = np.asarray(coord, np.float32)
    ndim = coord.shape[-1]
    oversamp = np.asarray([2.0] * ndim)
    shape = np.asarray([shape] * ndim if np.isscalar(shape) else list(shape)[-ndim:], np.int16)
    dcf = np.ones(coord.shape[:-1], np.float32) if dcf is None else np.asarray(dcf, np.float32)
    isreal = False
    if basis is not None:
        basis = np.asarray(basis)
        isreal = not np.iscomplexobj(basis)
        basis = basis.astype(np.float32 if isreal else np.complex64)
        adj = basis.conj().T  # (nframes, ncoeff)
        ncoeff = basis.shape[0]
        delta = np.ones((ncoeff,) + coord.shape[:-1], np.float32)
        delta = (delta.T * adj).T
    else:
        delta = np.ones(coord.shape[:-1], np.float32)
    is_cart = cartesian_axes(coord, shape)
    oversamp[is_cart] = 1.0
    kshape = np.ceil(oversamp * shape).astype(np.int16)
    kcoord = (coord * oversamp.astype(np.float32)).astype(np.float32)
    psf = nufft_adj((dcf * delta).astype(np.complex64), kcoord, [int(s) for s in kshape], width=width,
                    basis=None if basis is None else basis.astype(np.complex64))
    ax = tuple(range(-ndim, 0))
    kern = np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(psf, axes=ax), axes=ax, norm="ortho"), axes=ax)
    if kern.shape[0] == 1:
        kern = kern[0]
    if isreal:
        kern = kern.real
    kern = np.fft.ifftshift(kern, axes=ax)
    if basis is not None:
        kern = kern.reshape(*kern.shape[:2], -1).transpose(2, 1, 0)
    kern = kern / np.mean(np.abs(kern[kern != 0]))
    kern = np.nan_to_num(kern)
    return np.ascontiguousarray(kern).astype(np.float32 if isreal else np.complex64), \
        tuple(int(s) for s in kshape)


def apply_selfadj(image, kern, kshape, lowrank=False, matvec=False):
    """Pad -> FFT (ortho) -> kernel product -> IFFT (ortho) -> crop (dm/fft/nufft.py:609-669).

    Low rank (dm/fft/_interp/toeplitz.py:51-84, :109-130): k-space as
    (voxels, batch, ncoeff); deepmr's per-voxel product is
    out[j] += K[i][j] * in[j] over i (each coefficient scaled by the column
    sum of the voxel's kernel). ``matvec`` instead forms out[i] = sum_j
    K[i][j] * in[j], the Gram matrix product.
    """
    x = np.asarray(image).astype(np.complex64)
    ndim = len(kshape)
    oshape = x.shape[-ndim:]
    x = resize(x, list(x.shape[:-ndim]) + list(kshape))
    ax = tuple(range(-ndim, 0))
    k = np.fft.fftn(x, axes=ax, norm="ortho")
    if lowrank:
        shp = k.shape
        kin = k.reshape(-1, shp[-ndim - 1], int(np.prod(shp[-ndim:]))).transpose(2, 0, 1)
        if matvec:
            kout = np.einsum("vij,vbj->vbi", kern, kin)
        else:
            kout = kin * kern.sum(axis=1)[:, None, :]
        k = kout.transpose(1, 2, 0).reshape(shp)
    else:
        k = kern * k
    y = np.fft.ifftn(k, axes=ax, norm="ortho")
    return resize(y, list(y.shape[:-ndim]) + list(oshape)).astype(np.complex64)
