"""Non-Cartesian trajectories for the NUFFT (drop-in for deepmr.radial,
deepmr.radial_proj and deepmr.spiral, src/deepmr/_vobj/sampling/radial.py:17,
radial_proj.py:17, spiral.py:16).

The radial generators start from one full spoke through the k-space
centre, ``nsamples`` points evenly spaced from -N/2 to +N/2 (deepmr designs
it in rad/m, src/deepmr/_design/grad/noncartesian/radial.py:216-236,
normalises it to [-0.5, 0.5] with scale_traj and multiplies by N), and
rotate copies of it:

- ``radial``: in-plane rotations about z by n * dphi, dphi = (1 - 233/377) *
  360 degrees (the pseudo golden angle of period 377), acquisition order
  contrasts first, then views (radial.py:143-160);
- ``radial_proj``: the spoke is rotated in-plane about z by i * 360/n0
  degrees and the plane about x by j * dphi (or (i + j) * dphi for the
  shuffled "ga-sh" order) (radial_proj.py:152-186);
- ``spiral``: one interleaf k(s) = kr(s) exp(i theta(s)) with kr evenly
  spaced to kmax and d theta = d kr * FOV(kr) / nintl (an Archimedean
  spiral whose field of view may step up at the calibration / navigator radii,
  interpolated monotonically), src/deepmr/_design/grad/noncartesian/
  spiral.py:272-312, rotated like ``radial``. Without gradient limits
  (deepmr's defaults, gmax = smax = None) no gradient design reshapes it.

Density weights follow deepmr: |k| of the base spoke in rad/m for the 2D
radial trajectory, and for the projections the same times the distance from
each projection's rotation axis, normalised to unit sum
(src/deepmr/_design/grad/utils/dcf.py:20-73); Voronoi areas of the spiral's
nintl uniformly rotated interleaves (dcf.py:75-160). Readout timing (``t``,
``TE``) comes from deepmr's gradient designer and is not restated here.
"""

__all__ = ["Header", "radial", "radial_proj", "spiral"]

import math
from dataclasses import dataclass, field

import numpy as np
import torch

_DPHI = (1.0 - 233.0 / 377.0) * 360.0  # pseudo golden angle (degrees), period 377


@dataclass
class Header:
    """Acquisition header of a designed trajectory (the fields the NUFFT
    path reads from deepmr's Header): ``shape`` (nz, ny, nx), ``traj``
    (ncontrasts, nviews, nsamples, ndim) float32 in [-N/2, N/2] with x
    first, ``dcf`` density weights."""

    shape: torch.Tensor
    traj: torch.Tensor
    dcf: torch.Tensor
    user: dict = field(default_factory=dict)


def _expand_shape(shape):
    shape = [shape, 1] if np.isscalar(shape) else list(shape)
    while len(shape) < 3:
        shape = shape + [1]
    return [int(s) for s in shape]


def _base_spoke(mtx, osf=1.0, variant="fullspoke"):
    """deepmr's design spoke (radial.py:215-236): samples of |k| <= kmax in
    rad/m for a 1 mm isotropic grid, returned as (k in rad/m, k scaled to
    [-mtx/2, mtx/2])."""
    if variant not in ("fullspoke", "center-out"):
        raise ValueError(f"unknown radial variant {variant!r} (fullspoke, center-out)")
    res = 1e-3  # fov = mtx mm at 1 mm resolution
    kmax = math.pi / res
    nsamp = int(osf * mtx)  # Nyquist radial samples (nyquist.py:34-51)
    if variant == "fullspoke":
        k = np.linspace(0.0, 2.0, nsamp) - 1
    else:
        k = np.linspace(0.0, 1.0, nsamp // 2)
    k = k * kmax
    scaled = k / np.abs(k).max() / 2.0 * mtx
    return k, scaled


def _rot_z(phi):
    c, s = np.cos(phi), np.sin(phi)
    z, o = np.zeros_like(phi), np.ones_like(phi)
    return np.stack([np.stack([c, -s, z], -1), np.stack([s, c, z], -1), np.stack([z, z, o], -1)], 1)


def _rot_x(theta):
    c, s = np.cos(theta), np.sin(theta)
    z, o = np.zeros_like(theta), np.ones_like(theta)
    return np.stack([np.stack([o, z, z], -1), np.stack([z, c, -s], -1), np.stack([z, s, c], -1)], 1)


def radial(shape, nviews=None, variant="fullspoke"):
    """2D radial trajectory (deepmr.radial, src/deepmr/_vobj/sampling/radial.py:17-174).

    shape: in-plane matrix N, or (N, ncontrasts, nechoes). nviews defaults to
    int(pi * N) with one contrast, else 1. Returns a Header with traj
    (ncontrasts * nechoes, nviews, nsamples, 2).
    """
    shape = _expand_shape(shape)
    mtx, ncontrasts, nechoes = shape[0], shape[1], shape[2]
    if nviews is None:
        nviews = int(math.pi * mtx) if ncontrasts == 1 else 1
    k, spoke = _base_spoke(mtx, variant=variant)
    phi = np.deg2rad(np.arange(ncontrasts * nviews) * _DPHI)
    base = np.stack([spoke, 0 * spoke, 0 * spoke], 0)  # (3, nsamples), along x
    traj = np.einsum("bij,jn->bni", _rot_z(phi), base)[..., :2]  # (nrot, nsamples, 2)
    traj = traj.reshape(nviews, ncontrasts, -1, 2).swapaxes(0, 1)
    traj = np.repeat(traj, nechoes, axis=0)
    dcf = np.abs(k)  # analytical weight of the base spoke (dcf.py:20-31)
    return Header(shape=torch.as_tensor([mtx, mtx]),
                  traj=torch.as_tensor(np.ascontiguousarray(traj), dtype=torch.float32),
                  dcf=torch.as_tensor(dcf, dtype=torch.float32))


def radial_proj(shape, nviews=None, order="ga", variant="fullspoke"):
    """3D radial projections (deepmr.radial_proj, src/deepmr/_vobj/sampling/radial_proj.py:17-218).

    shape: N or (N, ncontrasts, nechoes). nviews: (in-plane spokes, planes);
    a scalar sets the planes (in-plane spokes int(pi * N)); the default is
    int(pi * N) planes with one contrast, else 1. order: "ga", "ga-sh"
    (shuffled), optionally with "::multiaxis" (views repeated on the three
    axis permutations). Returns a Header with traj
    (ncontrasts * nechoes, nviews, nsamples, 3).
    """
    shape = _expand_shape(shape)
    mtx, ncontrasts, nechoes = shape[0], shape[1], shape[2]
    if nviews is None:
        nviews = int(math.pi * mtx) if ncontrasts == 1 else 1
    if np.isscalar(nviews):
        nviews = [int(math.pi * mtx), int(nviews)]
    n0, n1 = int(nviews[0]), int(nviews[1])
    k, spoke = _base_spoke(mtx, variant=variant)
    j = np.tile(np.arange(ncontrasts * n1), n0)
    i = np.repeat(np.arange(n0), ncontrasts * n1)
    theta = (i + j) * _DPHI if order[:5] == "ga-sh" else j * _DPHI
    phi = i * (360.0 / n0)
    rot = np.einsum("bij,bjk->bik", _rot_x(np.deg2rad(theta)), _rot_z(np.deg2rad(phi)))
    base = np.stack([spoke, 0 * spoke, 0 * spoke], 0)
    traj = np.einsum("bij,jn->bni", rot, base)  # (nrot, nsamples, 3)
    traj = traj.reshape(n0, n1, ncontrasts, -1, 3).transpose(2, 1, 0, 3, 4)
    traj = traj.reshape(ncontrasts, -1, traj.shape[-2], 3)
    # in-plane |k| times the distance from the x rotation axis, normalised
    # (angular_compensation, dcf.py:34-73)
    flat = traj.reshape(-1, traj.shape[-2], 3)
    dcf = np.abs(k)[None, :] * np.sqrt((flat[..., 1:] ** 2).sum(-1))
    dcf = (dcf / dcf.sum()).reshape(traj.shape[:-1])
    if order[-9:] == "multiaxis":
        t1 = np.stack((traj[..., 2], traj[..., 0], traj[..., 1]), -1)
        t2 = np.stack((traj[..., 1], traj[..., 2], traj[..., 0]), -1)
        traj = np.concatenate((traj, t1, t2), axis=-3)
        dcf = np.concatenate((dcf, dcf, dcf), axis=-2)
        k0 = np.argmin((traj[0, 0] ** 2).sum(-1) ** 0.5)
        dcf = dcf / (dcf[..., [k0]] + 1e-6) / (n0 * n1 * ncontrasts)
    traj = np.repeat(traj, nechoes, axis=0)
    dcf = np.repeat(dcf, nechoes, axis=0)
    return Header(shape=torch.as_tensor([mtx, mtx, mtx]),
                  traj=torch.as_tensor(np.ascontiguousarray(traj), dtype=torch.float32),
                  dcf=torch.as_tensor(np.ascontiguousarray(dcf), dtype=torch.float32))


def _spiral_interleaf(mtx, nintl, acs_shape=None, acs_nintl=1, moco_shape=None, trans_dur=None,
                      variant="center-out"):
    """deepmr's spiral interleaf (spiral.py:172-312): returns (k (2, nsamples)
    in rad/m, nintl)."""
    from scipy import interpolate

    if variant not in ("center-out", "reverse", "in-out"):
        raise ValueError(f"unknown spiral variant {variant!r} (center-out, reverse, in-out)")
    mtxs = [moco_shape, acs_shape, mtx]
    nint = [1, acs_nintl, int(nintl)]
    nint = np.array([nint[n] for n in range(3) if mtxs[n] is not None])
    mtxs = np.array([m for m in mtxs if m is not None], dtype=np.float64)
    fov = mtx * 1e-3  # 1 mm isotropic
    res = fov / mtxs
    nr = int(math.ceil(fov / res.min()))  # Nyquist radial samples (nyquist.py:34-51)
    ntheta = int(math.ceil(0.5 * math.pi * math.ceil(fov / res.min())))
    if variant == "in-out":
        nint = nint * 2
    fovs = fov / (nint / nint.max())
    n_int = int(nint.max())
    nsamp = nr * ntheta // n_int
    kmax = math.pi / res.min()
    kr = kmax * np.linspace(0.0, 1.0, nsamp)
    dkr = np.diff(kr)[0]
    r = np.flip(res[1:] / res[0])
    if len(r) == 0:
        radius = np.asarray([0.0, 1.0])
    else:
        if trans_dur is None:
            trans_dur = r[0]
        if r.max() + trans_dur >= 1.0:
            raise ValueError("the transition band must end before kmax")
        r = np.repeat(r, 2)
        r[1::2] += trans_dur
        radius = np.concatenate(([0.0], r, [1.0]))
    fov_r = interpolate.PchipInterpolator(radius * kmax, np.repeat(fovs, 2))(kr)
    theta = np.cumsum(dkr * fov_r / n_int)
    k = kr * np.exp(1j * theta)
    k = np.stack([k.real, k.imag], 0)
    if variant == "reverse":
        k = np.flip(k, axis=-1)
    elif variant == "in-out":
        k = np.concatenate((-np.flip(k), k), axis=-1)
    return np.ascontiguousarray(k), (n_int // 2 if variant == "in-out" else n_int)


def spiral(shape, accel=None, nintl=1, acs_shape=None, acs_nintl=1, moco_shape=None, trans_dur=None,
           variant="center-out", dcf=True):
    """2D spiral (deepmr.spiral, src/deepmr/_vobj/sampling/spiral.py:16-201).

    shape: N or (N, ncontrasts, nechoes); nintl interleaves fully sample the
    plane, ``accel`` keeps nintl // accel of them (default 1 with one
    contrast, else nintl: one interleaf per contrast). ``acs_shape`` /
    ``moco_shape`` make the dual / triple density variants. Returns a Header
    with traj (ncontrasts * nechoes, nviews, nsamples, 2); ``dcf`` (Voronoi
    weights of the interleaf among its nintl rotations) unless dcf=False.
    """
    from . import dcf as _dcf

    shape = _expand_shape(shape)
    mtx, ncontrasts, nechoes = shape[0], shape[1], shape[2]
    if accel is None:
        accel = 1 if ncontrasts == 1 else nintl
    if not 0 < accel <= nintl:
        raise ValueError(f"acceleration {accel} must be in [1, {nintl}]")
    k, n_int = _spiral_interleaf(mtx, nintl, acs_shape, acs_nintl, moco_shape, trans_dur, variant)
    nviews = max(int(nintl // accel), 1)
    scaled = k / np.sqrt((k ** 2).sum(0)).max() / 2.0 * mtx
    phi = np.deg2rad(np.arange(ncontrasts * nviews) * _DPHI)
    base = np.stack([scaled[0], scaled[1], 0 * scaled[0]], 0)
    traj = np.einsum("bij,jn->bni", _rot_z(phi), base)[..., :2]
    traj = traj.reshape(nviews, ncontrasts, -1, 2).swapaxes(0, 1)
    traj = np.repeat(traj, nechoes, axis=0)
    w = _dcf.voronoi(k.T, n_int) if dcf else None
    user = {"moco_shape": moco_shape, "acs_shape": acs_shape}
    return Header(shape=torch.as_tensor([mtx, mtx]),
                  traj=torch.as_tensor(np.ascontiguousarray(traj), dtype=torch.float32),
                  dcf=None if w is None else torch.as_tensor(np.asarray(w), dtype=torch.float32), user=user)
