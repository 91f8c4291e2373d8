"""GPU parity of the non-Cartesian calibration caller ``find_acs``
(src/deepmr/recon/calib/acs.py:115-155): samples within the calibration
radius, density weighted, gridded by the HIP adjoint NUFFT (oversampling 2,
width 4), then the centred FFT over the last three axes.

Checked against the oracle's adjoint NUFFT run through the same selection
(relative L2 <= 1e-5, complex64) and against the FP64 direct adjoint DFT
divided by prod(os_shape) (deepmr's nufft_adj scaling), which fixes the
selection, weighting and axis order independently of the Kaiser-Bessel
approximation (relative L2 <= 5e-3: the KB error at width 4, osf 2 is ~1e-3)."""

import numpy as np
import pytest

from oracle import ndft
from oracle import nufft_oracle as orc
from tilenufft import calib

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def cfft3(x):
    ax = (-3, -2, -1)
    return np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(x, axes=ax), axes=ax, norm="ortho"), axes=ax)


def cnormal(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def test_find_acs_2d_radial():
    print("find_acs 2D radial", flush=True)
    rng = np.random.default_rng(11)
    nv, ns, cal = 24, 64, (16, 16)
    phi = np.arange(nv) * np.pi * (np.sqrt(5) - 1) / 2
    r = np.arange(ns) - ns / 2
    coord = np.stack([np.cos(phi)[:, None] * r, np.sin(phi)[:, None] * r], -1)[None].astype(np.float32)
    k = cnormal(rng, (2, 3, 1, nv, ns))  # (nslices, ncoils, ncontrasts, nviews, nsamples)
    w = (np.abs(r) + 0.5)[None, None, :].repeat(nv, 1).astype(np.float32)
    out = calib.find_acs(k, cal, coord, w, device=DEV).numpy()
    assert out.shape == (3, 2, 16, 16)  # (ncoils, nslices, ny, nx)

    sel = np.abs(r) <= 8
    ck = (w[..., sel] * k[..., sel]).astype(np.complex64)
    ref = orc.nufft_adj(ck, coord[..., sel, :], cal, oversamp=2.0)[:, :, 0].swapaxes(0, 1)
    assert rel(out, cfft3(ref)) < 1e-5
    direct = ndft.ndft_adjoint(ck.reshape(2, 3, -1), coord[..., sel, :], cal) / (32 * 32)
    assert rel(out, cfft3(direct.swapaxes(0, 1))) < 5e-3


def test_find_acs_3d_kooshball():
    print("find_acs 3D kooshball", flush=True)
    rng = np.random.default_rng(12)
    nv, ns, cal = 60, 16, (8, 8, 8)
    u = rng.standard_normal((nv, 3))
    u /= np.linalg.norm(u, axis=-1, keepdims=True)
    r = np.arange(ns) - ns / 2
    coord = (u[:, None, :] * r[None, :, None])[None].astype(np.float32)
    k = cnormal(rng, (4, 1, nv, ns))  # (ncoils, ncontrasts, nviews, nsamples)
    out = calib.find_acs(k, cal, coord, device=DEV).numpy()
    assert out.shape == (4, 8, 8, 8)
    sel = np.abs(r) <= 4
    ck = k[..., sel]
    ref = orc.nufft_adj(ck.reshape(4, 1, 1, nv, -1), coord[..., sel, :], cal, oversamp=2.0)[:, 0, 0]
    assert rel(out, cfft3(ref)) < 1e-5
    direct = ndft.ndft_adjoint(ck.reshape(4, -1), coord[..., sel, :], cal) / 16 ** 3
    assert rel(out, cfft3(direct)) < 5e-3
