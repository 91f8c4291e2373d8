"""GPU parity of the pruned, fused FFT stages of the NUFFT (tn_fft_forward /
tn_fft_adjoint; in 3D with the z stage tn_fft_z_forward / tn_fft_z_adjoint
around them) against the oracle's own steps (src/deepmr/fft/nufft.py:499-505
and :587-593): apodize -> centred zero pad -> centred FFT, and centred
inverse FFT -> centred crop -> apodize. The device grid holds the centred
spectrum with the shifts folded in (ifftshift of the oracle's), so the
forward stage is compared with ifftshift(oracle) and the adjoint is fed
fftshift of the device grid. Bar: relative L2 <= 1e-5 (complex64)."""

import numpy as np
import pytest
import torch

import tilenufft
from oracle import nufft_oracle as orc
from tilenufft import testing

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
RTOL = 1e-5


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-30))


def radial_coord(shape, nviews=8, nsamp=64, cart_y=False, seed=0):
    rng = np.random.default_rng(seed)
    ny, nx = shape
    c = np.stack([rng.uniform(-nx / 2, nx / 2, (nviews, nsamp)),
                  rng.uniform(-ny / 2, ny / 2, (nviews, nsamp))], -1).astype(np.float32)
    if cart_y:  # integer y positions: a Cartesian axis (width 1, no oversampling)
        c[..., 1] = np.floor(c[..., 1])
    return c


# (image shape (y, x), oversampling (scalar or (y, x)), Cartesian y)
CASES = [
    ((64, 64), 1.25, False),          # 80 = 8 * 2 * 5
    ((48, 64), 2.0, False),           # 96 x 128
    ((80, 96), 2.0, False),           # 160 x 192, radix 5 and 3
    ((51, 77), (1.88, 2.07), False),  # odd image sizes: 96 x 160
    ((256, 256), 2.0, False),         # 512 x 512
    ((320, 160), 2.0, False),         # 640 x 320
    ((512, 384), (2.0, 2.0), False),  # 1024 x 768
    ((64, 48), 2.0, True),            # Cartesian y: 64 (no padding) x 96
    ((192, 192), 1.25, False),        # 240 = 8 * 2 * 3 * 5
    ((320, 320), 1.25, False),        # 400 = 8 * 2 * 5 * 5
    ((384, 192), (1.5, 1.25), False),  # 576 x 240
]


@pytest.mark.parametrize("shape, osf, cart_y", CASES)
def test_fft_stages_match_oracle(shape, osf, cart_y):
    print(f"fft stages shape={shape} osf={osf} cart_y={cart_y}", flush=True)
    coord = radial_coord(shape, cart_y=cart_y)
    plan = tilenufft.fft.plan_nufft(coord, shape, 4, osf, device=DEV)
    native = plan.interpolator
    assert native.has_fft, plan.os_shape
    if cart_y:
        assert plan.width[0] == 1 and plan.os_shape[0] == shape[0]
    ref = orc.plan_nufft(coord, shape, 4, osf)
    assert tuple(ref["os_shape"]) == tuple(plan.os_shape)
    nimg = 3
    x = testing.complex_normal((nimg,) + shape, seed=1).numpy()
    g = torch.empty((nimg,) + tuple(plan.os_shape), dtype=torch.complex64, device=DEV)
    native.fft_forward(torch.from_numpy(x).to(DEV), g, nimg)
    xa = x.copy()
    orc.apodize(xa, 2, ref["oversamp"], ref["width"], ref["beta"])
    k_ref = orc.fft_c(orc.resize(xa, [nimg] + list(ref["os_shape"])), (-2, -1))
    k_ref = np.fft.ifftshift(k_ref, axes=(-2, -1))
    assert rel(g.cpu().numpy(), k_ref) < RTOL

    grid = testing.complex_normal((nimg,) + tuple(plan.os_shape), seed=2).numpy()
    img = torch.empty((nimg,) + shape, dtype=torch.complex64, device=DEV)
    native.fft_adjoint(torch.from_numpy(grid).to(DEV), img, nimg)
    y = orc.ifft_c(np.fft.fftshift(grid, axes=(-2, -1)), (-2, -1))
    y = orc.resize(y, [nimg] + list(shape)).astype(np.complex64)
    orc.apodize(y, 2, ref["oversamp"], ref["width"], ref["beta"])
    assert rel(img.cpu().numpy(), y) < RTOL


def test_fft_stages_end_to_end_with_plain_path():
    """apply_nufft / apply_nufft_adj on a plan with the pruned FFT stages agree
    with the oracle at the workload's grid (256^2 image, 512^2 grid)."""
    n = 256
    print("fft stages end to end", flush=True)
    coord = testing.vd_spiral(6, 512, n)
    plan = tilenufft.fft.plan_nufft(coord, n, 4, 2.0, device=DEV)
    assert plan.interpolator.has_fft
    x = testing.complex_normal((2, n, n), seed=3).numpy()
    y = tilenufft.fft.apply_nufft(torch.from_numpy(x).to(DEV), plan).cpu().numpy()
    p_ref = orc.plan_nufft(coord, n, 4, 2.0)
    assert rel(y, orc.apply_nufft(x, p_ref)) < RTOL
    k = testing.complex_normal(y.shape, seed=4).numpy()
    xa = tilenufft.fft.apply_nufft_adj(torch.from_numpy(k).to(DEV), plan).cpu().numpy()
    assert rel(xa, orc.apply_nufft_adj(k, p_ref)) < RTOL


def test_fft_stages_unsupported_lengths_fall_back():
    """A grid length with a prime factor above 5 (here 7 * 8 = 56) has no
    pruned stages; the NUFFT then runs the separate pad / crop kernels around
    the library FFT and still matches the oracle."""
    print("fft stages fallback", flush=True)
    coord = radial_coord((28, 28))
    plan = tilenufft.fft.plan_nufft(coord, 28, 4, 2.0, device=DEV)
    assert tuple(plan.os_shape) == (56, 56) and not plan.interpolator.has_fft
    with pytest.raises(RuntimeError, match="pruned FFT"):
        plan.interpolator.fft_forward(torch.zeros((1, 28, 28), dtype=torch.complex64, device=DEV),
                                      torch.zeros((1, 56, 56), dtype=torch.complex64, device=DEV), 1)
    x = testing.complex_normal((1, 28, 28), seed=5).numpy()
    y = tilenufft.fft.apply_nufft(torch.from_numpy(x).to(DEV), plan).cpu().numpy()
    assert rel(y, orc.apply_nufft(x, orc.plan_nufft(coord, 28, 4, 2.0))) < RTOL


@pytest.mark.parametrize("shape, osf", [((32, 32, 32), 2.0), ((64, 64, 64), 1.25), ((64, 64, 64), 1.5),
                                         ((51, 51, 51), 1.25), ((16, 32, 48), 2.0), ((40, 24, 32), 2.0)])
def test_fft_stages_3d_match_library_fft(shape, osf):
    """3D plans: the native z stage around the (y, x) plane stages against
    tn_apod_pad + torch.fft.fftn (forward) and torch.fft.ifftn +
    tn_crop_apod (adjoint), for even and odd image sizes (N = 51: grid 64,
    the centring of an odd axis) and for three different axis lengths
    (z, y, x), so that an exchanged axis or pitch shows."""
    nz, ny, nx = shape
    coord = testing.kooshball(20, 16, 2) * (np.array([nx, ny, nz], np.float32) / 2)
    plan = tilenufft.fft.plan_nufft(coord, shape, 4, osf, device=DEV)
    nat = plan.interpolator
    assert nat.has_fft
    nimg = 3
    g = plan.os_shape
    x = testing.complex_normal((nimg, *shape), seed=nx).to(DEV)
    grid = torch.empty((nimg, *g), dtype=torch.complex64, device=DEV)
    got = nat.fft_forward(x, grid, nimg).clone()
    pad = torch.empty_like(grid)
    nat.apod_pad(x, pad, nimg)
    ref = torch.fft.fftn(pad, dim=(-3, -2, -1))
    assert rel(got.cpu().numpy(), ref.cpu().numpy()) < 1e-5
    k = testing.complex_normal((nimg, *g), seed=nx + 1).to(DEV)
    img = torch.empty((nimg, *shape), dtype=torch.complex64, device=DEV)
    ref_img = torch.empty_like(img)
    nat.crop_apod(torch.fft.ifftn(k, dim=(-3, -2, -1)), ref_img, nimg)
    nat.fft_adjoint(k.clone(), img, nimg)
    assert rel(img.cpu().numpy(), ref_img.cpu().numpy()) < 1e-5


def test_stack_of_stars_short_z_axis_matches_oracle():
    """Stack of stars: golden-angle radial in-plane, integer z (a Cartesian
    slice axis: width 1, osf 1, so Gz = Nz = 32, a short transform length).
    The forward and adjoint run through the 3D FFT stages and match the
    oracle."""
    n, nz = 32, 32
    print(f"stack of stars {nz} slices {n}^2", flush=True)
    rad = testing.golden_angle_radial(12, 48, n)  # (views, samples, 2)
    z = (np.arange(nz) - nz // 2).astype(np.float32)
    coord = np.concatenate([np.broadcast_to(rad[None], (nz, *rad.shape)),
                            np.broadcast_to(z[:, None, None, None], (nz, *rad.shape[:2], 1))], -1)
    coord = np.ascontiguousarray(coord.reshape(nz * rad.shape[0], rad.shape[1], 3))
    plan = tilenufft.fft.plan_nufft(coord, (nz, n, n), 4, 2.0, device=DEV)
    assert plan.os_shape == (nz, 2 * n, 2 * n) and plan.width[0] == 1
    assert plan.interpolator.has_fft
    x = testing.complex_normal((2, nz, n, n), seed=40).numpy()
    y = tilenufft.fft.apply_nufft(torch.from_numpy(x).to(DEV), plan).cpu().numpy()
    p_ref = orc.plan_nufft(coord, (nz, n, n), 4, 2.0)
    assert rel(y, orc.apply_nufft(x, p_ref)) < 1e-5
    k = testing.complex_normal(y.shape, seed=41).numpy()
    xa = tilenufft.fft.apply_nufft_adj(torch.from_numpy(k).to(DEV), plan).cpu().numpy()
    assert rel(xa, orc.apply_nufft_adj(k, p_ref)) < 1e-5


def test_anisotropic_3d_nufft_through_fft_stages_matches_oracle():
    """A 3D image with three different axis lengths (z, y, x) = (8, 16, 24)
    at osf 2: grid (16, 32, 48), every axis through the native stages. The
    forward and adjoint NUFFT match the oracle."""
    shape = (8, 16, 24)
    print(f"anisotropic 3D NUFFT {shape}", flush=True)
    coord = testing.kooshball(40, 24, 2) * (np.array([24, 16, 8], np.float32) / 2)
    plan = tilenufft.fft.plan_nufft(coord, shape, 4, 2.0, device=DEV)
    assert plan.os_shape == (16, 32, 48) and plan.interpolator.has_fft
    x = testing.complex_normal((2, *shape), seed=42).numpy()
    y = tilenufft.fft.apply_nufft(torch.from_numpy(x).to(DEV), plan).cpu().numpy()
    p_ref = orc.plan_nufft(coord, shape, 4, 2.0)
    assert rel(y, orc.apply_nufft(x, p_ref)) < RTOL
    k = testing.complex_normal(y.shape, seed=43).numpy()
    xa = tilenufft.fft.apply_nufft_adj(torch.from_numpy(k).to(DEV), plan).cpu().numpy()
    assert rel(xa, orc.apply_nufft_adj(k, p_ref)) < RTOL
