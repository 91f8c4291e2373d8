"""CPU restatement of deepmr's Cartesian sparse FFT (test infrastructure).

Only the tests, the smoke check and the bench's CPU baseline use this module;
it is the checker, never part of the product path.

Follows src/deepmr/fft/sparse_fft.py:21-503 and src/deepmr/fft/_sparse/*.py:
  * sampling mask: integer k-space indexes (ncontrasts, nviews, nsamples,
    ndim), x first, on the centred grid of shape (z, y, x)
    (_sparse/plan.py:14-94);
  * sparse FFT: centred FFT with norm=None, then a gather at the indexes
    (sparse_fft.py:331-395, _sparse/dense2sparse.py:120-198);
  * sparse IFFT: scatter-add of the samples (duplicates accumulate) into a
    zero grid, then the centred inverse FFT (1/prod(shape) scaling)
    (sparse_fft.py:398-456, _sparse/sparse2dense.py:177-239);
  * low-rank variants contract basis_adjoint (nframes, ncoeff) in the gather
    and basis (ncoeff, nframes) in the scatter (dense2sparse.py:201-290,
    sparse2dense.py:243-329);
  * Toeplitz: the kernel is the scattered mask itself (ifftshifted,
    normalised by the mean non-zero magnitude), applied between uncentred
    orthonormal FFTs (_sparse/toeplitz.py:18-114, sparse_fft.py:459-503).
Layout rules: nframes == 1 gives batch = all leading dims, else the dims
before the contrast axis; outputs swap back (dense2sparse.py:84-118,
sparse2dense.py:84-125).

Pins: the reference's own known answers (tests/fft/test_sparse_fft.py:
centred delta <-> all ones, self-adjoint delta -> delta, identity bases),
the FP64 direct DFT at the integer frequencies and the dot-product identity.
"""

import numpy as np

from .nufft_oracle import fft_c, ifft_c


def plan_sampling(indexes, shape):
    """Integer indexes (..., ndim), x first -> dict with per-axis (z, y, x)
    index arrays of shape (nframes, npts) (_sparse/plan.py:56-94)."""
    idx = np.asarray(indexes)
    ndim = idx.shape[-1]
    ishape = list(idx.shape[:-1])
    while len(ishape) < 3:
        ishape = [1] + ishape
    nframes = int(ishape[0])
    ishape = tuple(int(s) for s in ishape[1:])
    shape = [int(shape)] * ndim if np.isscalar(shape) else [int(s) for s in list(shape)[-ndim:]]
    idx = idx.astype(np.int16).astype(np.int64).reshape(nframes, int(np.prod(ishape)), ndim)
    index = tuple(idx[..., d] for d in range(ndim))[::-1]  # (z, y, x)
    for d in range(ndim):
        if index[d].size and (index[d].min() < 0 or index[d].max() >= shape[d]):
            raise ValueError("sampling indexes outside the grid")
    return dict(index=index, dshape=tuple(shape), ishape=ishape, ndim=ndim, nframes=nframes)


def _flat(mask):
    """Flat grid offsets (nframes, npts) of the sampled points."""
    off = np.zeros(mask["index"][0].shape, np.int64)
    for d, n in enumerate(mask["dshape"]):
        off = off * n + mask["index"][d]
    return off


def sample_raw(cart, mask, basis_adjoint=None):
    """(nframes|ncoeff, batch, *dshape) -> (nframes, batch, npts)."""
    off = _flat(mask)
    nframes = off.shape[0]
    cart = np.asarray(cart, np.complex64)
    flat = cart.reshape(cart.shape[0], cart.shape[1], -1)
    if basis_adjoint is None:
        return np.stack([flat[f][:, off[f]] for f in range(nframes)]).astype(np.complex64)
    ba = np.asarray(basis_adjoint, np.complex64)
    out = np.zeros((nframes, cart.shape[1], off.shape[1]), np.complex64)
    for f in range(nframes):
        for c in range(flat.shape[0]):
            out[f] += ba[f, c] * flat[c][:, off[f]]
    return out


def zerofill_raw(noncart, mask, basis=None):
    """(nframes, batch, npts) -> (nframes|ncoeff, batch, *dshape), duplicates add."""
    off = _flat(mask)
    nframes, batch, _ = noncart.shape
    vol = int(np.prod(mask["dshape"]))
    noncart = np.asarray(noncart, np.complex64)
    if basis is None:
        out = np.zeros((nframes, batch, vol), np.complex64)
        for f in range(nframes):
            for b in range(batch):
                np.add.at(out[f, b], off[f], noncart[f, b])
    else:
        bs = np.asarray(basis, np.complex64)
        out = np.zeros((bs.shape[0], batch, vol), np.complex64)
        for c in range(bs.shape[0]):
            for f in range(nframes):
                for b in range(batch):
                    np.add.at(out[c, b], off[f], bs[c, f] * noncart[f, b])
    return out.reshape(out.shape[0], batch, *mask["dshape"])


def apply_sampling(data_in, mask, basis_adjoint=None):
    """Layout rules of _sparse/dense2sparse.py:14-118."""
    x = np.asarray(data_in).astype(np.complex64)
    ndim, nframes, dshape, ishape = mask["ndim"], mask["nframes"], mask["dshape"], mask["ishape"]
    batch_shape = x.shape[:-ndim] if nframes == 1 else x.shape[: -ndim - 1]
    batch = int(np.prod(batch_shape))
    lead = nframes if basis_adjoint is None else np.asarray(basis_adjoint).shape[-1]
    x = x.reshape(batch, lead, *dshape).swapaxes(0, 1)
    out = sample_raw(x, mask, basis_adjoint)
    if nframes == 1:
        return out[0].reshape(*batch_shape, *ishape)
    return out.swapaxes(0, 1).reshape(*batch_shape, nframes, *ishape)


def apply_zerofill(data_in, mask, basis=None):
    """Layout rules of _sparse/sparse2dense.py:14-125."""
    y = np.asarray(data_in).astype(np.complex64)
    nframes, dshape = mask["nframes"], mask["dshape"]
    npts = int(np.prod(mask["ishape"]))
    batch_shape = y.shape[:-2] if nframes == 1 else y.shape[:-3]
    batch = int(np.prod(batch_shape))
    y = y.reshape(batch, nframes, npts).swapaxes(0, 1)
    out = zerofill_raw(y, mask, basis)
    if nframes == 1:
        return out[0].reshape(*batch_shape, *dshape)
    return out.swapaxes(0, 1).reshape(*batch_shape, out.shape[0], *dshape)


def apply_sparse_fft(image, mask, basis_adjoint=None, weight=None):
    """sparse_fft.py:331-395."""
    ndim = mask["ndim"]
    k = fft_c(np.asarray(image).astype(np.complex64), range(-ndim, 0)).astype(np.complex64)
    y = apply_sampling(k, mask, basis_adjoint)
    if weight is not None:
        y = (np.asarray(weight, np.float32) * y).astype(np.complex64)
    return y


def apply_sparse_ifft(kspace, mask, basis=None, weight=None):
    """sparse_fft.py:398-456."""
    ndim = mask["ndim"]
    y = np.asarray(kspace).astype(np.complex64)
    if weight is not None:
        y = (np.asarray(weight, np.float32) * y).astype(np.complex64)
    g = apply_zerofill(y, mask, basis)
    return ifft_c(g, range(-ndim, 0)).astype(np.complex64)


def sparse_fft(image, indexes, basis_adjoint=None):
    """deepmr.fft.sparse_fft (src/deepmr/fft/__init__.py:33-86)."""
    ndim = np.asarray(indexes).shape[-1]
    mask = plan_sampling(indexes, np.asarray(image).shape[-ndim:])
    return apply_sparse_fft(image, mask, basis_adjoint)


def sparse_ifft(kspace, indexes, shape, basis=None):
    """deepmr.fft.sparse_ifft (src/deepmr/fft/__init__.py:89-142)."""
    return apply_sparse_ifft(kspace, plan_sampling(indexes, shape), basis)


def plan_toeplitz_fft(indexes, shape, basis=None):
    """_sparse/toeplitz.py:18-114 -> (kernel, islowrank)."""
    idx = np.asarray(indexes)
    ndim = idx.shape[-1]
    isreal = False
    if basis is not None:
        basis = np.asarray(basis)
        isreal = not np.iscomplexobj(basis)
        basis = basis.astype(np.float32 if isreal else np.complex64)
        ncoeff = basis.shape[0]
        delta = np.ones((ncoeff,) + idx.shape[:-1], np.float32)
        delta = (delta.T * basis.conj().T).T
    else:
        delta = np.ones(idx.shape[:-1], np.float32)
    mask = plan_sampling(idx, shape)
    kern = apply_zerofill(delta.astype(np.complex64), mask, None if basis is None else basis)
    if isreal:
        kern = kern.real
    ax = tuple(range(-ndim, 0))
    kern = np.fft.ifftshift(kern, axes=ax)
    if basis is not None:
        kern = kern.reshape(*kern.shape[:2], -1).transpose(2, 1, 0)
    kern = kern / np.mean(np.abs(kern[kern != 0]))
    kern = np.nan_to_num(kern)
    return np.ascontiguousarray(kern).astype(np.float32 if isreal else np.complex64), basis is not None


def apply_sparse_fft_selfadj(image, kern, ndim, lowrank=False, matvec=False):
    """sparse_fft.py:459-503 with _sparse/toeplitz.py:117-183 (deepmr's
    column-sum product for low rank, or the Gram mat-vec with ``matvec``)."""
    x = np.asarray(image).astype(np.complex64)
    ax = tuple(range(-ndim, 0))
    k = np.fft.fftn(x, axes=ax, norm="ortho")
    if lowrank:
        shp = k.shape
        kin = k.reshape(-1, shp[-ndim - 1], int(np.prod(shp[-ndim:]))).transpose(2, 0, 1)
        if matvec:
            kout = np.einsum("vij,vbj->vbi", kern, kin)
        else:
            kout = kin * kern.sum(axis=1)[:, None, :]
        k = kout.transpose(1, 2, 0).reshape(shp)
    else:
        k = kern * k
    return np.fft.ifftn(k, axes=ax, norm="ortho").astype(np.complex64)
