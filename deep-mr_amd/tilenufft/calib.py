"""k-space calibration region for coil-sensitivity estimation
(deepmr: src/deepmr/recon/calib/acs.py:1-155).

``find_acs`` is the calibration caller of the adjoint NUFFT: for
non-Cartesian data it keeps the samples within the calibration radius,
grids them with ``nufft_adj`` (oversampling 2, width 4) on the HIP device and
returns the centred FFT of the low-resolution coil images. The Cartesian
branch is host code (the largest fully sampled block around the centre) and
runs where the data lives, as deepmr's does.
"""

__all__ = ["find_acs"]

import numpy as np
import torch

from . import fft as _fft


def find_acs(data, shape=None, coord=None, dcf=None, device=None):
    """Extract the k-space calibration region (acs.py:11-48).

    data layouts (deepmr's): 2D Cartesian (nslices, ncoils, ncontrasts, ny, nx);
    2D non-Cartesian (nslices, ncoils, ncontrasts, nviews, nsamples);
    3D Cartesian (nx, ncoils, ncontrasts, nz, ny); 3D non-Cartesian
    (ncoils, ncontrasts, nviews, nsamples).
    shape: calibration matrix (non-Cartesian only), e.g. (24, 24).
    coord: (ncontrasts, nviews, nsamples, ndim) in [-N/2, N/2), x first;
    None selects the Cartesian branch. dcf: (ncontrasts, nviews, nsamples).
    device: HIP device for the adjoint NUFFT (default: the data's, or the
    current one).
    Returns complex64 calibration k-space (ncoils, nz, ny, nx), a torch tensor.
    """
    if coord is None:
        return _find_cart_acs(data)
    return _find_noncart_acs(data, shape, coord, dcf, device)


def _largest_block(mask):
    """Bounds [l, r) per axis of the fully sampled block around the centre:
    grow a cube from the centre voxel while every face is sampled, then
    stretch each axis left and right in turn (acs.py:74-97)."""
    sl = [[n // 2, n // 2 + 1] for n in mask.shape]

    def full(bounds):
        return bool(np.all(mask[tuple(slice(l, r) for l, r in bounds)]))

    while all(l > 0 and r < mask.shape[i] for i, (l, r) in enumerate(sl)) and \
            full([[l - 1, r + 1] for l, r in sl]):
        sl = [[l - 1, r + 1] for l, r in sl]
    for d in range(mask.ndim):
        while sl[d][0] > 0 and full([[l - (k == d), r] for k, (l, r) in enumerate(sl)]):
            sl[d][0] -= 1
        while sl[d][1] < mask.shape[d] and full([[l, r + (k == d)] for k, (l, r) in enumerate(sl)]):
            sl[d][1] += 1
    return sl


def _find_cart_acs(kspace):
    """Cartesian branch (acs.py:52-112): sum the contrasts, find the block
    sampled in coil 0, crop every coil to it, put coils first and take the
    centred FFT along the slice (or readout) axis."""
    istorch = isinstance(kspace, torch.Tensor)
    device = kspace.device if istorch else None
    k = kspace.sum(axis=-3)
    k = k.numpy(force=True) if istorch else np.asarray(k)
    k = np.moveaxis(k.astype(np.complex64), 1, -1)  # coils last
    sl = _largest_block(np.abs(k[..., 0]) > 0)
    k = np.moveaxis(k[tuple(slice(l, r) for l, r in sl) + (slice(None),)].copy(), -1, 1)
    k = torch.as_tensor(k, dtype=torch.complex64, device=device) if istorch else k
    k = k.swapaxes(0, 1)  # (ncoils, nz, ny, nx)
    return _fft.fft(k, axes=(1,))


def _find_noncart_acs(kspace, cal_shape, coord, dcf, device):
    """Non-Cartesian branch (acs.py:115-155)."""
    ndim = coord.shape[-1]
    kspace = kspace.reshape(*kspace.shape[:2], 1, -1, kspace.shape[-1])
    coord = coord.reshape(1, -1, *coord.shape[-2:])
    if dcf is not None:
        dcf = dcf.reshape(1, -1, dcf.shape[-1])

    # samples of the first view within the calibration radius select the
    # readout positions kept on every view
    cabs = (coord ** 2).sum(axis=-1) ** 0.5
    ind = cabs[0, 0] <= min(cal_shape) // 2
    if isinstance(ind, torch.Tensor):
        ind = ind.cpu()
    ind = np.asarray(ind)
    cal_ksp = kspace[..., ind]
    cal_coord = coord[..., ind, :]
    if dcf is not None:
        cal_ksp = dcf[..., ind] * cal_ksp

    if device is None and isinstance(kspace, torch.Tensor) and kspace.device.type != "cpu":
        device = kspace.device
    cal = _fft.nufft_adj(cal_ksp, cal_coord, cal_shape, oversamp=2.0,
                         device=device if device is not None else "cpu")
    cal = torch.as_tensor(cal).squeeze()
    if cal.ndim != 4:
        cal = cal[None, ...]
    if ndim == 2:
        cal = cal.swapaxes(0, 1)
    cal = _fft.fft(cal, axes=range(-3, 0))
    return torch.as_tensor(cal, dtype=torch.complex64)
