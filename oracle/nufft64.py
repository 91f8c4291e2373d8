"""ORACLE — test infrastructure only.

Float64 restatement of deepmr's Kaiser-Bessel NUFFT (same steps as
oracle/nufft_oracle.py, every quantity in float64 / complex128), for the
finite-difference gradient checks that deepmr's own operator cannot serve:
deepmr casts every input to complex64 (src/deepmr/fft/nufft.py:461-464).

Steps (src/deepmr/fft/nufft.py:450-606): apodize, centred zero-pad, centred
FFT (norm=None), Kaiser-Bessel interpolation / gridding with tap origin
ceil(g - W/2) mod G (src/deepmr/fft/_interp/plan.py:251-269), 1/scale; the
adjoint ends with the centred IFFT (x 1/prod(G)), crop and apodization.
"""

import math

import numpy as np

from . import nufft_oracle as orc


def plan(coord, shape, width=4, oversamp=1.25):
    """Geometry from the float32 oracle plan; taps and apodization in float64."""
    p = orc.plan_nufft(coord, shape, width, oversamp)
    sc = p["scaled_coord"].astype(np.float64)
    ndim = p["ndim"]
    ishape = list(sc.shape[:-1])
    while len(ishape) < 3:
        ishape = [1] + ishape
    nframes = ishape[0]
    npts = int(np.prod(ishape[1:]))
    sc = sc.reshape(nframes, npts, ndim)
    idx, val = [], []  # (z, y, x) order, each (nframes, npts, W)
    for ax in range(ndim):  # zyx axis index
        d = ndim - 1 - ax  # coordinate column
        w, g_size, beta = p["width"][ax], p["os_shape"][ax], float(p["beta"][ax])
        g = sc[..., d]
        if w == 1:
            idx.append(g.astype(np.int64)[..., None])
            val.append(np.ones(g.shape + (1,)))
            continue
        x0 = np.ceil(g - w / 2)
        k = np.arange(w)
        pos = x0[..., None] + k
        idx.append(np.mod(pos, g_size).astype(np.int64))
        val.append(orc.kaiser_bessel((pos - g[..., None]) / (w / 2), beta).reshape(pos.shape))
    scale = orc.kernel_scaling(p["beta"], p["width"])
    apod = []
    for ax in range(ndim):
        n, w = p["shape"][ax], p["width"][ax]
        if w == 1:
            apod.append(np.ones(n))
            continue
        gsz = math.ceil(float(np.float32(p["oversamp"][ax])) * n)
        i = np.arange(n) - n // 2
        s = np.sqrt((float(p["beta"][ax]) ** 2 - (math.pi * w * i / gsz) ** 2).astype(complex))
        a = np.real(s / np.sinh(s))
        a = a / a[n // 2]
        apod.append(np.nan_to_num(a, nan=1.0))
    return dict(p, idx=idx, val=val, scale=scale, apod=apod, nframes=nframes, npts=npts,
                ishape=tuple(ishape[1:]))


def _apod(x, pl):
    ndim = pl["ndim"]
    for ax in range(ndim):
        shp = [1] * x.ndim
        shp[x.ndim - ndim + ax] = -1
        x = x * pl["apod"][ax].reshape(shp)
    return x


def _taps(pl):
    """Yield (weight, index tuple) per tap over the W^d footprint."""
    ndim = pl["ndim"]
    ranges = [range(v.shape[-1]) for v in pl["val"]]
    for ks in np.ndindex(*[len(r) for r in ranges]):
        w = np.ones(pl["val"][0].shape[:-1])
        index = []
        for ax in range(ndim):
            w = w * pl["val"][ax][..., ks[ax]]
            index.append(pl["idx"][ax][..., ks[ax]])
        yield w, index


def forward(x, pl):
    """(..., [nframes], *N) -> (..., [nframes], *ishape) in complex128."""
    ndim, nf = pl["ndim"], pl["nframes"]
    x = _apod(np.asarray(x, np.complex128), pl)
    x = orc.resize(x, list(x.shape[:-ndim]) + list(pl["os_shape"]))
    ax = tuple(range(-ndim, 0))
    k = np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(x, axes=ax), axes=ax), axes=ax)
    lead = k.shape[:-ndim] if nf == 1 else k.shape[:-ndim - 1]
    kk = k.reshape(-1, nf, *pl["os_shape"])
    out = np.zeros((kk.shape[0], nf, pl["npts"]), np.complex128)
    frames = np.arange(nf)[:, None]
    for w, index in _taps(pl):
        out += w[None] * kk[(slice(None), frames) + tuple(index)]
    out /= pl["scale"]
    shape = (*lead, *pl["ishape"]) if nf == 1 else (*lead, nf, *pl["ishape"])
    return out.reshape(shape)


def adjoint(y, pl):
    """deepmr's adjoint (A^H / prod(os_shape)) in complex128."""
    ndim, nf = pl["ndim"], pl["nframes"]
    y = np.asarray(y, np.complex128)
    lead = y.shape[:-2] if nf == 1 else y.shape[:-3]
    yy = y.reshape(-1, nf, pl["npts"])
    g = np.zeros((yy.shape[0], nf, *pl["os_shape"]), np.complex128)
    frames = np.broadcast_to(np.arange(nf)[:, None], (nf, pl["npts"]))
    for w, index in _taps(pl):
        for b in range(yy.shape[0]):
            np.add.at(g[b], (frames,) + tuple(np.broadcast_to(i, (nf, pl["npts"])) for i in index),
                      w * yy[b])
    g /= pl["scale"]
    ax = tuple(range(-ndim, 0))
    x = np.fft.fftshift(np.fft.ifftn(np.fft.ifftshift(g, axes=ax), axes=ax), axes=ax)
    x = orc.resize(x, list(x.shape[:-ndim]) + list(pl["shape"]))
    x = _apod(x, pl)
    shape = (*lead, *pl["shape"]) if nf == 1 else (*lead, nf, *pl["shape"])
    return x.reshape(shape)
