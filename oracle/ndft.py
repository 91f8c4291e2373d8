"""ORACLE — test infrastructure only.

FP64 direct non-uniform DFT: the closed form that deepmr's Kaiser-Bessel NUFFT
approximates. With image index n_d in [0, N_d) and r_d = n_d - N_d//2 (the
centred convention of dm/fft/fft.py:76-80 after the centred zero-pad of
dm/_signal/resize.py:86-97), and coordinates c_{k,d} in grid units
(coord[..., 0] = x, the last image axis):

    forward   y_k = sum_r x_r * exp(-2*pi*i * sum_d c_{k,d} r_d / N_d)
    adjoint   x_r = sum_k y_k * exp(+2*pi*i * sum_d c_{k,d} r_d / N_d)

deepmr's nufft_adj equals the adjoint above divided by prod(os_shape)
(norm=None on both FFTs, dm/fft/nufft.py:505, :587).
"""

import numpy as np


def _phase_matrix(coord, shape_zyx, sign):
    coord = np.asarray(coord, np.float64).reshape(-1, coord.shape[-1])
    ndim = coord.shape[-1]
    shape_xyz = list(shape_zyx)[::-1]
    axes = [np.arange(n) - n // 2 for n in shape_zyx]  # (z, y, x)
    grids = np.meshgrid(*axes, indexing="ij")
    r_xyz = [g.ravel() for g in grids[::-1]]  # x first
    arg = np.zeros((coord.shape[0], r_xyz[0].size))
    for d in range(ndim):
        arg += np.outer(coord[:, d] / shape_xyz[d], r_xyz[d])
    return np.exp(sign * 2j * np.pi * arg)


def ndft(image, coord, shape_zyx=None):
    """Forward NDFT of image (..., *shape) at coord (..., ndim) -> (..., npts)."""
    coord = np.asarray(coord)
    ndim = coord.shape[-1]
    image = np.asarray(image, np.complex128)
    if shape_zyx is None:
        shape_zyx = image.shape[-ndim:]
    E = _phase_matrix(coord, shape_zyx, -1.0)
    flat = image.reshape(-1, int(np.prod(shape_zyx)))
    return (flat @ E.T).reshape(*image.shape[:-ndim], coord.reshape(-1, ndim).shape[0])


def ndft_adjoint(kspace, coord, shape_zyx):
    """Adjoint NDFT of kspace (..., npts) -> (..., *shape)."""
    coord = np.asarray(coord)
    E = _phase_matrix(coord, shape_zyx, -1.0)
    flat = np.asarray(kspace, np.complex128).reshape(-1, E.shape[0])
    return (flat @ E.conj()).reshape(*np.asarray(kspace).shape[:-1], *shape_zyx)
