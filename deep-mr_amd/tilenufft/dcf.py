"""Density compensation factors for non-Cartesian trajectories, the weights
callers pass to ``apply_nufft_adj`` / ``NUFFTOp``
(deepmr: src/deepmr/_design/grad/utils/dcf.py:1-363).

Host code (numpy + scipy.spatial.Voronoi), as in deepmr: these run once per
trajectory, at design time, not per transform.

Behaviour kept from deepmr:
  * ``voronoi`` trims the largest cell volumes at the value of rank
    round((1 - 0.9 * 0.98) * M) in descending order (dcf.py:228-239), leaves
    open (infinite) cells at 0 unless ``fix_edge`` (dcf.py:242-250), then
    inverts twice with zero guards (``_normalize_weight``, dcf.py:162-169) and
    scales so the sample nearest the k-space centre weighs 1 / nshots.
  * 3D cell volumes use deepmr's fan formula over the region's vertex order
    (dcf.py:300-314): sum |(B x C) . A^T| / 6 over all pairs. It equals the
    true volume for a tetrahedron, not for a general polyhedron.
Departures: torch input is accepted (deepmr's branch calls ``.numpy``
without parentheses, dcf.py:95, and cannot run); repeated k-space
locations share the first occurrence's weight evenly, as deepmr's docstring
states (dcf.py:80-81).
"""

__all__ = ["analytical_dcf", "angular_compensation", "voronoi"]

import numpy as np
import torch


def analytical_dcf(coord):
    """|k|: the radius of each sample (dcf.py:20-31)."""
    return (coord ** 2).sum(axis=-1) ** 0.5


def angular_compensation(wi, coord, rotation_axis):
    """In-plane weights ``wi`` (npts,) times the radius of each projection
    with its rotation axis dropped, normalised to unit sum (dcf.py:34-72).

    coord: (nproj, npts, 3); rotation_axis: (nproj,) in {0, 1, 2}."""
    coord = np.asarray(coord)
    keep = [np.delete(coord[n], int(rotation_axis[n]), axis=-1) for n in range(len(rotation_axis))]
    w = np.asarray(wi)[None, :] * analytical_dcf(np.stack(keep, 0))
    return w / np.sum(w)


def _uniform_rotations(nshots):
    """In-plane rotations by 2*pi*n/nshots about z (deepmr's "uniform" tilt,
    tilt.py:14-54 with tilt_increment "uniform" = 2*pi/N)."""
    phi = (np.arange(nshots) * (2.0 * np.pi / nshots)) % (2.0 * np.pi)
    c, s = np.cos(phi), np.sin(phi)
    return c, s


def voronoi(coord, nshots=1, fix_edge=False, ratio=0.1):
    """Voronoi-cell density compensation (dcf.py:75-158).

    coord: (..., 2) or (..., 3) real coordinates, or complex (kx + i ky).
    With ``nshots`` > 1 the single interleaf ``coord`` (npts, 2) is rotated
    uniformly to a full disk, the cells are computed on all shots, and the
    first shot's weights are returned. The centre sample is found on the
    flattened samples (deepmr indexes the reshaped weights with a flat index,
    which only works for (M, ndim) input, dcf.py:125-155)."""
    istorch = isinstance(coord, torch.Tensor)
    device = coord.device if istorch else None
    coord = coord.detach().cpu().numpy().copy() if istorch else np.array(coord, copy=True)
    if np.iscomplexobj(coord):
        coord = np.stack((coord.real, coord.imag), axis=-1)
    coord = coord.astype(np.float32)

    if nshots > 1:
        c, s = _uniform_rotations(nshots)
        x, y = coord[..., 0], coord[..., 1]
        coord = np.stack((c[:, None] * x[None] - s[:, None] * y[None],
                          s[:, None] * x[None] + c[:, None] * y[None]), -1).astype(np.float32)
        cabs = (coord[0] ** 2).sum(axis=-1) ** 0.5
    else:
        cabs = (coord ** 2).sum(axis=-1) ** 0.5
    k0_idx = np.argmin(cabs)

    ishape = coord.shape[:-1]
    flat = coord.reshape(-1, coord.shape[-1])
    _, first, inverse, count = np.unique(flat, return_index=True, return_inverse=True,
                                         return_counts=True, axis=0)
    inverse = inverse.ravel()
    if np.any(count > 1):
        wi = np.zeros(flat.shape[0], dtype=np.float32)
        wi[first] = _voronoi(flat[first], fix_edge, ratio)
        share = wi[first][inverse] / count[inverse]  # split each location's weight evenly
        wi = share.astype(np.float32)
    else:
        wi = _voronoi(flat, fix_edge, ratio)
    wi = np.nan_to_num(wi, posinf=0.0, neginf=0.0).reshape(ishape)
    if nshots > 1:
        wi = wi[0]
    wi = _normalize_weight(wi)
    wi = (1.0 / wi.reshape(-1)[k0_idx] / nshots) * wi
    wi = wi.squeeze()
    if istorch:
        wi = torch.as_tensor(wi, dtype=torch.float32, device=device)
    return wi


def _normalize_weight(weights):
    """Double inversion with zero guards (dcf.py:162-169)."""
    weights = np.array(weights, copy=True)
    total = np.sum(weights)
    weights[np.isclose(weights, 0.0)] = 1.0
    inv = total / weights
    inv = inv / np.sum(inv)
    inv[np.isclose(inv, 0.0)] = 1.0
    return 1.0 / inv


def polygon_area(points):
    """Shoelace area of a polygon given by its vertices in order (dcf.py:267-286)."""
    p = np.asarray(points, np.float64)
    x, y = p[:, 0], p[:, 1]
    return abs(np.sum(x * (np.roll(y, -1) - np.roll(y, 1)))) / 2.0


def fan_volume(points):
    """deepmr's 3D cell measure (dcf.py:300-314): with p0 the first vertex,
    A = p[:-2] - p0, B = p[1:-1] - p0, C = p[2:] - p0, sum |(B x C) A^T| / 6."""
    p = np.asarray(points, np.float64)
    a, b, c = p[:-2] - p[0], p[1:-1] - p[0], p[2:] - p[0]
    return np.sum(np.abs(np.cross(b, c) @ a.T)) / 6.0


def _voronoi(coord, fix_edge, ratio):
    """Cell volumes of unique points, trimmed (dcf.py:206-252)."""
    from scipy.spatial import Voronoi

    vol = polygon_area if coord.shape[-1] == 2 else fan_volume
    vor = Voronoi(coord)
    wi = np.zeros(coord.shape[0], dtype=np.float32)
    for m in range(coord.shape[0]):
        region = vor.regions[vor.point_region[m]]
        if len(region) and all(v != -1 for v in region):
            wi[m] = vol(vor.vertices[np.asarray(region, np.int64)])

    fufa = 0.98
    while True:
        r = 1 - (1 - ratio) * fufa
        order = np.argsort(wi)[::-1]
        thld = wi[order[int(round(r * wi.size))]]
        if np.isfinite(thld):
            break
        fufa **= 2
    wi[wi > thld] = thld

    if fix_edge:  # extrapolate open cells from the outer 40% of radii^2 (cubic fit)
        rho = np.sum(coord.astype(np.float64) ** 2, axis=-1)
        good = rho > 0.6 * np.max(rho)
        poly = np.polynomial.Polynomial.fit(rho[good], wi[good], 3)
        wi[wi == 0] = poly(rho[wi == 0])
    return wi
