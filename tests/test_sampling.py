"""Trajectory generators (CPU only): deepmr.radial / deepmr.radial_proj.

- the reference's own shape known answers (tests/vobj/test_sampling.py:118-160);
- every spoke runs from -N/2 to +N/2 through the centre with evenly spaced
  samples, consecutive 2D views are rotated by the pseudo golden angle
  (1 - 233/377) * 360 degrees, and contrasts are acquired before views;
- 3D projections: in-plane angles i * 360 / n0, plane tilts about x by the
  pseudo golden angle; the density weights are |k| times the distance from
  the x axis, normalised to unit sum;
- the trajectories feed the NUFFT plan unchanged (coordinates inside
  [-N/2, N/2], no Cartesian axis).
"""

import math

import numpy as np
import pytest

from tilenufft import sampling

DPHI = np.deg2rad((1 - 233 / 377) * 360.0)


@pytest.mark.parametrize("args, kw, shape", [
    ((128,), {}, (1, 402, 128, 2)),
    ((128,), {"nviews": 64}, (1, 64, 128, 2)),
    (((128, 420),), {}, (420, 1, 128, 2)),
    (((128, 1, 8),), {}, (8, 402, 128, 2)),
])
def test_radial_shapes_match_reference(args, kw, shape):
    head = sampling.radial(*args, **kw)
    assert tuple(head.traj.shape) == shape
    assert tuple(head.dcf.shape) == (shape[2],)


@pytest.mark.parametrize("args, kw, shape", [
    ((128,), {}, (1, 161604, 128, 3)),
    ((128,), {"nviews": 64}, (1, 25728, 128, 3)),
    ((128,), {"nviews": (64, 402)}, (1, 25728, 128, 3)),
    (((128, 420),), {}, (420, 402, 128, 3)),
    (((128, 1, 8),), {}, (8, 161604, 128, 3)),
])
def test_radial_proj_shapes_match_reference(args, kw, shape):
    head = sampling.radial_proj(*args, **kw)
    assert tuple(head.traj.shape) == shape
    assert tuple(head.dcf.shape) == shape[:3]


def test_radial_geometry():
    n = 64
    t = sampling.radial(n, nviews=10).traj.numpy()[0]  # (views, samples, 2)
    r = np.linalg.norm(t, axis=-1)
    assert np.allclose(r.max(axis=1), n / 2, atol=1e-4)
    # evenly spaced samples along a line through the centre
    d = np.diff(t, axis=1)
    assert np.allclose(d, d[:, :1], atol=1e-4)
    assert np.allclose(np.linalg.norm(d, axis=-1), n / (n - 1), atol=1e-4)
    # view v points along angle v * dphi (the spoke runs from -k to +k)
    ang = np.arctan2(t[:, -1, 1], t[:, -1, 0])
    assert np.allclose(np.exp(1j * ang), np.exp(1j * DPHI * np.arange(10)), atol=1e-5)
    # contrasts first: with 3 contrasts x 2 views, contrast c of view v has angle (v * 3 + c) * dphi
    t3 = sampling.radial((n, 3), nviews=2).traj.numpy()
    a3 = np.arctan2(t3[:, :, -1, 1], t3[:, :, -1, 0])
    want = DPHI * (np.arange(2)[None, :] * 3 + np.arange(3)[:, None])
    assert np.allclose(np.exp(1j * a3), np.exp(1j * want), atol=1e-5)
    dcf = sampling.radial(n).dcf.numpy()
    assert np.allclose(dcf / dcf.max(), np.abs(np.linspace(-1, 1, n)), atol=1e-6)


def test_radial_proj_geometry_and_weights():
    n, n0, n1 = 32, 6, 5
    h = sampling.radial_proj(n, nviews=(n0, n1))
    t = h.traj.numpy()[0].reshape(n1, n0, n, 3)  # (plane j, in-plane i, samples, xyz)
    end = t[..., -1, :] / (n / 2)  # unit direction of each spoke
    assert np.allclose(np.linalg.norm(end, axis=-1), 1, atol=1e-5)
    i = np.arange(n0)[None, :]
    j = np.arange(n1)[:, None]
    phi = np.deg2rad(i * 360.0 / n0)
    th = DPHI * j
    want = np.stack([np.cos(phi) + 0 * th, np.sin(phi) * np.cos(th), np.sin(phi) * np.sin(th)], -1)
    assert np.allclose(end, want, atol=1e-5)
    w = h.dcf.numpy()
    assert abs(w.sum() - 1) < 1e-5
    k = np.abs(np.linspace(-1, 1, n))
    flat = h.traj.numpy()[0]
    expect = k[None, :] * np.linalg.norm(flat[..., 1:], axis=-1)
    assert np.allclose(w[0], expect / expect.sum(), rtol=1e-5, atol=1e-12)


def test_trajectories_plan_as_noncartesian():
    from tilenufft import _core

    t2 = sampling.radial(32, nviews=12).traj.numpy()
    t3 = sampling.radial_proj(16, nviews=(8, 6)).traj.numpy()
    for t, n in ((t2, 32), (t3, 16)):
        assert np.abs(t).max() <= n / 2 + 1e-4
        assert not _core.cartesian_axes(t, [n] * t.shape[-1]).any()


@pytest.mark.parametrize("args, kw, shape", [
    ((128,), {}, (1, 1, 25856, 2)),
    ((128,), {"nintl": 48}, (1, 48, 538, 2)),
    ((128,), {"nintl": 48, "accel": 4}, (1, 12, 538, 2)),
    (((128, 420),), {"nintl": 48}, (420, 1, 538, 2)),
    (((128, 1, 8),), {"nintl": 48}, (8, 48, 538, 2)),
])
def test_spiral_shapes_match_reference(args, kw, shape):
    """deepmr's tests/vobj/test_sampling.py:36-55."""
    head = sampling.spiral(*args, dcf=False, **kw)
    assert tuple(head.traj.shape) == shape


def test_spiral_geometry_and_density():
    """An Archimedean interleaf: radius in grid units grows by nintl per turn,
    so nintl rotated copies are one grid unit apart (Nyquist); it spans the
    radius N/2; the dual-density variant turns nintl / acs_nintl times faster
    inside the calibration radius; Voronoi weights grow with the radius."""
    n, nintl = 64, 8
    h = sampling.spiral(n, nintl=nintl)
    t = h.traj.numpy()[0, 0].astype(np.float64)
    rho = np.linalg.norm(t, axis=-1)
    theta = np.unwrap(np.arctan2(t[:, 1], t[:, 0]))
    assert abs(rho.max() - n / 2) < 1e-4
    sel = rho > 2
    slope = np.polyfit(rho[sel], theta[sel], 1)[0]
    assert abs(slope - 2 * np.pi / nintl) < 1e-3 * 2 * np.pi / nintl
    w = h.dcf.numpy()
    assert w.shape == (t.shape[0],)
    outer = rho > n / 4
    assert np.median(w[outer]) > np.median(w[(rho > 2) & (rho < n / 8)])
    d = sampling.spiral(n, nintl=nintl, acs_shape=16, acs_nintl=2, dcf=False).traj.numpy()[0, 0].astype(np.float64)
    rd = np.linalg.norm(d, axis=-1)
    td = np.unwrap(np.arctan2(d[:, 1], d[:, 0]))
    inner = (rd > 0.5) & (rd < 0.25 * n / 2)
    s_in = np.polyfit(rd[inner], td[inner], 1)[0]
    assert abs(s_in - 2 * np.pi / nintl * nintl / 2) < 0.02 * s_in
