"""Writes the golden NUFFT vectors under tests/golden/ (run from the repo root).

The vectors are produced by the oracle (the CPU restatement of deepmr's
Kaiser-Bessel NUFFT, oracle/nufft_oracle.py), not by deepmr itself: deepmr
cannot be imported here (its kernels need Numba, which is absent). They pin
the oracle against regressions and give the GPU tests fixed expected outputs.

    python tests/golden/make_golden.py
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "deep-mr_amd"))

from oracle import nufft_oracle as orc  # noqa: E402
from tilenufft import testing  # noqa: E402


def cases():
    rng = np.random.default_rng(2024)
    # 2D: golden-angle radial, N=16 (config-1 geometry at small size)
    c2 = testing.golden_angle_radial(12, 32, 16)
    x2 = (rng.standard_normal((2, 16, 16)) + 1j * rng.standard_normal((2, 16, 16))).astype(np.complex64)
    k2 = (rng.standard_normal((2, 12, 32)) + 1j * rng.standard_normal((2, 12, 32))).astype(np.complex64)
    yield "nufft2d_radial_n16_osf2", c2, 16, 4, 2.0, x2, k2
    # 2D anisotropic, osf 1.25 (deepmr's default)
    c2b = np.stack([rng.uniform(-10, 10, (6, 20)), rng.uniform(-6, 6, (6, 20))], -1).astype(np.float32)
    x2b = (rng.standard_normal((12, 20)) + 1j * rng.standard_normal((12, 20))).astype(np.complex64)
    k2b = (rng.standard_normal((6, 20)) + 1j * rng.standard_normal((6, 20))).astype(np.complex64)
    yield "nufft2d_aniso_12x20_osf125", c2b, (12, 20), 4, 1.25, x2b, k2b
    # 3D kooshball N=8
    c3 = testing.kooshball(24, 16, 8)
    x3 = (rng.standard_normal((8, 8, 8)) + 1j * rng.standard_normal((8, 8, 8))).astype(np.complex64)
    k3 = (rng.standard_normal((24, 16)) + 1j * rng.standard_normal((24, 16))).astype(np.complex64)
    yield "nufft3d_koosh_n8_osf15", c3, 8, 4, 1.5, x3, k3


def main():
    for name, coord, shape, width, osf, x, k in cases():
        y = orc.nufft(x, coord, shape, width, osf)
        xa = orc.nufft_adj(k, coord, shape, width, osf)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), coord=coord,
                            shape=np.atleast_1d(np.asarray(shape)), width=width, oversamp=osf,
                            image=x, kspace=k, nufft=y, nufft_adj=xa)
        print(name, y.shape, xa.shape)


if __name__ == "__main__":
    main()
