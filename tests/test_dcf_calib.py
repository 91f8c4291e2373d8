"""Density compensation (src/deepmr/_design/grad/utils/dcf.py) and the
Cartesian calibration region (src/deepmr/recon/calib/acs.py:52-112): host
code, pinned by closed forms -- Voronoi cells of rectilinear and polar
lattices, shoelace / tetrahedron volumes, and a sampled block whose bounds are
known by construction."""

import numpy as np
import pytest
import torch

from tilenufft import calib, dcf


def test_polygon_area_and_tetrahedron_volume():
    square = [[0, 0], [2, 0], [2, 3], [0, 3]]
    assert dcf.polygon_area(square) == pytest.approx(6.0)
    hexagon = [[np.cos(a), np.sin(a)] for a in np.arange(6) * np.pi / 3]
    assert dcf.polygon_area(hexagon) == pytest.approx(3 * np.sqrt(3) / 2)
    p = np.array([[0.3, -0.2, 0.1], [2.0, 0.1, 0.0], [0.4, 1.7, 0.2], [0.1, 0.3, 1.4]])
    vol = abs(np.linalg.det(p[1:] - p[0])) / 6.0  # the tetrahedron's true volume
    assert dcf.fan_volume(p) == pytest.approx(vol)


def test_analytical_and_angular():
    c = np.array([[3.0, 4.0, 0.0], [0.0, 0.0, 2.0]])
    np.testing.assert_allclose(dcf.analytical_dcf(c), [5.0, 2.0])
    coord = np.array([[[3.0, 4.0, 12.0], [1.0, 0.0, 0.0]], [[3.0, 4.0, 12.0], [0.0, 2.0, 0.0]]])
    w = dcf.angular_compensation(np.array([1.0, 2.0]), coord, np.array([2, 0]))
    ang = np.array([[5.0, 1.0], [np.hypot(4, 12), 2.0]])  # radius with the rotation axis dropped
    ref = np.array([1.0, 2.0])[None] * ang
    np.testing.assert_allclose(w, ref / ref.sum())


def _expected(areas, k0, nshots=1):
    """Trimmed at rank round(0.118 M) (descending), zeros -> 1, scaled so the
    centre sample weighs 1 / nshots."""
    w = np.array(areas, np.float64)
    thld = np.sort(w)[::-1][int(round((1 - 0.9 * 0.98) * w.size))]
    w = np.minimum(w, thld)
    w[np.isclose(w, 0.0)] = 1.0
    return w / w[k0] / nshots


def test_voronoi_rectilinear_lattice():
    xs = np.cumsum([0, 1, 1, 1.5, 2, 1, 0.5, 1, 2, 1, 1])
    ys = np.cumsum([0, 0.5, 1, 1, 2, 1, 1, 3, 1])
    xs, ys = xs - xs[5], ys - ys[4]  # the centre sample sits at (0, 0)
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel()], -1)
    areas = np.zeros(pts.shape[0])
    for i in range(1, len(xs) - 1):
        for j in range(1, len(ys) - 1):
            areas[i * len(ys) + j] = (xs[i + 1] - xs[i - 1]) / 2 * (ys[j + 1] - ys[j - 1]) / 2
    k0 = 5 * len(ys) + 4
    w = dcf.voronoi(pts)
    np.testing.assert_allclose(w, _expected(areas, k0), rtol=1e-5)
    # complex input, torch input
    np.testing.assert_allclose(dcf.voronoi(pts[:, 0] + 1j * pts[:, 1]), w, rtol=1e-6)
    wt = dcf.voronoi(torch.from_numpy(pts.astype(np.float32)))
    assert isinstance(wt, torch.Tensor)
    np.testing.assert_allclose(wt.numpy(), w, rtol=1e-6)


def test_voronoi_duplicates_share_the_cell():
    xs = np.arange(9.0) - 4
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel()], -1)
    dup = np.concatenate([pts, pts[[40, 40, 30]]], 0)  # centre three times, (3, -1) twice
    areas = np.zeros(dup.shape[0])
    inner = (np.abs(dup[:, 0]) < 4) & (np.abs(dup[:, 1]) < 4)
    areas[inner] = 1.0
    for i in (40, 81, 82):
        areas[i] = 1.0 / 3
    for i in (30, 83):
        areas[i] = 0.5
    w = dcf.voronoi(dup)
    np.testing.assert_allclose(w, _expected(areas, 40), rtol=1e-5)


def test_voronoi_polar_interleaves():
    """One centre-out spoke rotated to 12 uniform shots: interior cells are
    trapezoids 2 r dr tan(pi/n), the shared centre a regular n-gon of
    apothem dr/2 split over the n copies, the outer ring open."""
    n, npts, dr = 12, 10, 1.0
    r = np.arange(npts) * dr
    spoke = np.stack([r, 0 * r], -1).astype(np.float32)
    areas = 2 * r * dr * np.tan(np.pi / n)
    areas[0] = n * (dr / 2) ** 2 * np.tan(np.pi / n) / n
    areas[-1] = 0.0
    full = np.concatenate([areas, np.tile(areas, n - 1)])  # the trimming rank runs over all shots
    thld = np.sort(full)[::-1][int(round((1 - 0.9 * 0.98) * full.size))]
    w = np.minimum(areas, thld)
    w[np.isclose(w, 0.0)] = 1.0
    np.testing.assert_allclose(dcf.voronoi(spoke, nshots=n), w / w[0] / n, rtol=1e-4)


def test_voronoi_fix_edge_fills_open_cells():
    xs = np.arange(11.0) - 5
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel()], -1)
    w = dcf.voronoi(pts, fix_edge=True)
    assert np.all(np.isfinite(w)) and np.all(w > 0)


def _cart_kspace(rng):
    nsl, nc, ncon, ny, nx = 3, 2, 2, 16, 20
    mask = np.zeros((ny, nx), bool)
    mask[5:11, 6:15] = True  # the calibration block around the centre (8, 10)
    mask[[0, 2, 14], :] = True  # isolated outer lines
    k = (rng.standard_normal((nsl, nc, ncon, ny, nx)) + 1j * rng.standard_normal((nsl, nc, ncon, ny, nx)))
    return (k * mask).astype(np.complex64)


@pytest.mark.parametrize("as_torch", [False, True])
def test_find_acs_cartesian_block(as_torch):
    k = _cart_kspace(np.random.default_rng(3))
    out = calib.find_acs(torch.from_numpy(k) if as_torch else k)
    out = out.numpy() if as_torch else out
    crop = k.sum(axis=-3)[:, :, 5:11, 6:15].swapaxes(0, 1)  # (coils, slices, 6, 9)
    ref = np.fft.fftshift(np.fft.fft(np.fft.ifftshift(crop, axes=1), axis=1, norm="ortho"), axes=1)
    assert out.shape == (2, 3, 6, 9)
    np.testing.assert_allclose(out, ref, rtol=1e-5, atol=1e-5)


def test_find_acs_cartesian_full_sampling_keeps_everything():
    k = np.ones((1, 1, 1, 8, 6), np.complex64)
    out = calib.find_acs(k)
    assert out.shape == (1, 1, 8, 6)
