"""Drop-in for the NUFFT and sparse-FFT parts of ``deepmr.fft``
(src/deepmr/fft/__init__.py:33-286)."""

__all__ = [
    "fft",
    "ifft",
    "nufft",
    "nufft_adj",
    "plan_nufft",
    "apply_nufft",
    "apply_nufft_adj",
    "NUFFTPlan",
    "set_autograd_mode",
    "get_autograd_mode",
    "plan_toeplitz_nufft",
    "apply_nufft_selfadj",
    "sparse_fft",
    "sparse_ifft",
    "prepare_sampling",
    "apply_sparse_fft",
    "apply_sparse_ifft",
    "plan_toeplitz_fft",
    "apply_sparse_fft_selfadj",
]

from .. import _interp  # noqa: F401  (deepmr.fft._interp)
from .fft import fft, ifft  # noqa: F401
from .toeplitz import apply_nufft_selfadj, plan_toeplitz_nufft  # noqa: F401
from .sparse_fft import (  # noqa: F401
    apply_sparse_fft,
    apply_sparse_fft_selfadj,
    apply_sparse_ifft,
    plan_toeplitz_fft,
    prepare_sampling,
)
from .nufft import (  # noqa: F401
    NUFFTPlan,
    apply_nufft,
    apply_nufft_adj,
    get_autograd_mode,
    plan_nufft,
    set_autograd_mode,
)


def sparse_fft(image, indexes, basis_adjoint=None, device="cpu", threadsperblock=128):
    """N-dimensional sparse FFT (src/deepmr/fft/__init__.py:33-86): centred
    FFT of ``image`` (..., [ncontrasts], *shape) sampled at the integer
    ``indexes`` (ncontrasts, nviews, nsamples, ndim), x first."""
    ndim = indexes.shape[-1]
    shape = image.shape[-ndim:]
    mask = prepare_sampling(indexes, shape, device)
    return apply_sparse_fft(image, mask, basis_adjoint, threadsperblock=threadsperblock)


def sparse_ifft(kspace, indexes, shape, basis=None, device="cpu", threadsperblock=128):
    """N-dimensional inverse sparse FFT (src/deepmr/fft/__init__.py:89-142)."""
    mask = prepare_sampling(indexes, shape, device)
    return apply_sparse_ifft(kspace, mask, basis, threadsperblock=threadsperblock)


def nufft(image, coord, shape=None, basis_adjoint=None, device="cpu", threadsperblock=128,
          width=4, oversamp=1.25):
    """N-dimensional NUFFT (deepmr ``nufft``, src/deepmr/fft/__init__.py:145-218).

    image (..., ncontrasts, ny, nx) or (..., ncontrasts, nz, ny, nx);
    coord (ncontrasts, nviews, nsamples, ndim) in [-N/2, N/2), x first.
    Returns k-space (..., ncontrasts, nviews, nsamples) on the image's device.
    Plans on every call, like deepmr; reuse ``plan_nufft`` + ``apply_nufft``
    to amortise planning.
    """
    ndim = coord.shape[-1]
    if shape is None:
        shape = image.shape[-ndim:]
    plan = plan_nufft(coord, shape, width, oversamp, device)
    return apply_nufft(image, plan, basis_adjoint, threadsperblock=threadsperblock)


def nufft_adj(kspace, coord, shape, basis=None, device="cpu", threadsperblock=128, width=4,
              oversamp=1.25):
    """Adjoint NUFFT (deepmr ``nufft_adj``, src/deepmr/fft/__init__.py:221-286).

    Same scaling as deepmr: the exact adjoint divided by prod(os_shape).
    """
    plan = plan_nufft(coord, shape, width, oversamp, device)
    return apply_nufft_adj(kspace, plan, basis, threadsperblock=threadsperblock)
