"""Synthetic inputs for tests and benchmarks: trajectories, coil maps, data.

All generators are deterministic (numpy/torch generators seeded by the
caller). Coordinates are returned in deepmr's convention: float32, last axis
(x, y[, z]) in grid units of the *image* matrix, i.e. in [-N/2, N/2).
"""

__all__ = [
    "GOLDEN_ANGLE",
    "golden_angle_radial",
    "vd_spiral",
    "kooshball",
    "gaussian_coil_maps",
    "complex_normal",
]

import math

import numpy as np
import torch

GOLDEN_ANGLE = math.pi * (math.sqrt(5.0) - 1.0) / 2.0  # 111.25 degrees
GOLDEN_MEANS = (0.4656, 0.6823)


def _radius(nsamples, n):
    return (np.arange(nsamples) - nsamples // 2) * (n / nsamples)


def golden_angle_radial(nspokes, nsamples, n, offset=0.0):
    """Golden-angle radial spokes through the centre, (nspokes, nsamples, 2)."""
    theta = offset + np.arange(nspokes) * GOLDEN_ANGLE
    r = _radius(nsamples, n)
    kx = np.outer(np.cos(theta), r)
    ky = np.outer(np.sin(theta), r)
    return np.stack([kx, ky], -1).astype(np.float32)


def vd_spiral(ninterleaves, nsamples, n, rotation=0.0, alpha=2.0):
    """Variable-density Archimedean spiral, (ninterleaves, nsamples, 2).

    k(tau) = kmax * tau^alpha * exp(i(2 pi turns tau + 2 pi l / L + rotation)),
    tau in [0, 1), kmax = N/2 - 1/2, turns = N / (2 L) * alpha: dense centre,
    Nyquist-ish periphery.
    """
    tau = np.arange(nsamples) / nsamples
    kmax = n / 2 - 0.5
    turns = n / (2.0 * ninterleaves) * alpha
    arms = []
    for l in range(ninterleaves):
        phase = 2 * np.pi * turns * tau + 2 * np.pi * l / ninterleaves + rotation
        k = kmax * tau**alpha * np.exp(1j * phase)
        arms.append(np.stack([k.real, k.imag], -1))
    return np.asarray(arms, dtype=np.float32)


def kooshball(nspokes, nsamples, n):
    """3D radial spokes on golden-means directions, (nspokes, nsamples, 3)."""
    s = np.arange(nspokes)
    kz = np.mod(s * GOLDEN_MEANS[0], 1.0) * 2.0 - 1.0
    phi = 2.0 * np.pi * np.mod(s * GOLDEN_MEANS[1], 1.0)
    st = np.sqrt(np.clip(1.0 - kz**2, 0.0, None))
    direction = np.stack([st * np.cos(phi), st * np.sin(phi), kz], -1)  # (x, y, z)
    r = _radius(nsamples, n)
    return (direction[:, None, :] * r[None, :, None]).astype(np.float32)


def gaussian_coil_maps(ncoils, shape, width=0.5, seed=0):
    """Smooth complex sensitivities: Gaussian bumps on a ring (2D) or sphere
    (3D) with linear phase, normalised so that sum_c |S_c|^2 = 1."""
    rng = np.random.default_rng(seed)
    ndim = len(shape)
    axes = [np.linspace(-1, 1, s, endpoint=False) + 1.0 / s for s in shape]
    grids = np.meshgrid(*axes, indexing="ij")  # (z, y, x) or (y, x)
    maps = []
    for c in range(ncoils):
        if ndim == 2:
            a = 2 * np.pi * c / ncoils
            centre = (0.9 * math.sin(a), 0.9 * math.cos(a))  # (y, x)
        else:
            z = 1 - 2 * (c + 0.5) / ncoils
            a = c * GOLDEN_ANGLE * 2
            rr = math.sqrt(max(0.0, 1 - z * z))
            centre = (0.9 * z, 0.9 * rr * math.sin(a), 0.9 * rr * math.cos(a))
        d2 = sum((g - c0) ** 2 for g, c0 in zip(grids, centre))
        mag = np.exp(-d2 / (2 * width**2))
        phase = sum(rng.uniform(-1, 1) * g for g in grids) + rng.uniform(-np.pi, np.pi)
        maps.append(mag * np.exp(1j * phase))
    maps = np.asarray(maps)
    maps /= np.sqrt((np.abs(maps) ** 2).sum(0, keepdims=True))
    return maps.astype(np.complex64)


def complex_normal(shape, generator=None, seed=None, device="cpu"):
    """Complex data with real and imaginary parts i.i.d. N(0, 1) (complex64)."""
    if generator is None:
        generator = torch.Generator().manual_seed(0 if seed is None else seed)
    re = torch.randn(shape, generator=generator, dtype=torch.float32)
    im = torch.randn(shape, generator=generator, dtype=torch.float32)
    return torch.complex(re, im).to(device)
