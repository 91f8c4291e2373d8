"""Non-Cartesian SENSE encoding and differentiable unrolled CG-SENSE.

``encoding_op`` builds E = F * C and E^H E like deepmr's EncodingOp NUFFT
branch (src/deepmr/recon/alg/linop.py:108-148): F is a NUFFTOp carrying the
square-rooted density weight, C a SenseOp. ``UnrolledCGSENSE`` runs
``niter`` CGStep iterations (src/deepmr/optim/cg.py:150-212) from zero on
E^H E x = E^H (dcf^1/2 y), as deepmr's recon_lstsq does with cg_solve
(src/deepmr/recon/alg/classic_recon.py:151-176) but with autograd enabled, so
gradients flow to the k-space (and to anything upstream of it) through the
hand-written forward / adjoint kernels.
"""

__all__ = ["encoding_op", "UnrolledCGSENSE", "cg_sense", "capture_unrolled"]

import torch
import torch.nn as nn

from . import linops, optim


def encoding_op(coord, shape, sensmap, dcf=None, device=None, width=4, oversamp=1.25,
                multicontrast=None):
    """Return (E, EHE) for non-Cartesian SENSE.

    coord  : (ncontrasts, nviews, nsamples, ndim) or (nviews, nsamples, ndim)
    shape  : image shape (..., ny, nx) / (..., nz, ny, nx)
    sensmap: (nsets, ncoils, *N) or (ncoils, *N) (a set axis is added)
    """
    ndim = coord.shape[-1]
    device = device if device is not None else (sensmap.device if isinstance(sensmap, torch.Tensor) else None)
    F = linops.NUFFTOp(coord, list(shape)[-ndim:], weight=dcf, device=device, width=width,
                       oversamp=oversamp)
    smap = torch.as_tensor(sensmap)
    if smap.dim() == ndim + 1:
        smap = smap[None]
    if multicontrast is None:
        multicontrast = len(coord.shape) >= 4
    C = linops.SenseOp(ndim, smap.to(F.nufft_plan.interpolator.device), multicontrast=multicontrast)
    E = F * C
    EHE = C.H * F.H * F * C
    return E, EHE


def cg_sense(kspace, E, EHE, niter=10, lamda=0.0, dcf=None):
    """Unrolled (differentiable) CG-SENSE: returns x after ``niter`` steps."""
    y = kspace if dcf is None else torch.as_tensor(dcf, device=kspace.device) ** 0.5 * kspace
    if _needs_set_axis(E, y):
        y = y[None]
    AHy = E.H(y)
    op = EHE if lamda == 0.0 else EHE + lamda * linops.Identity(EHE.ndim)
    step = optim.CGStep(op, AHy, ndim=EHE.ndim)
    x = torch.zeros_like(AHy)
    for _ in range(niter):
        x = step(x)
    return x


def _needs_set_axis(E, y):
    # SenseOp outputs (nsets, ..., ncoils, ...): add the set axis when missing
    sense = E.linops[-1]
    return y.dim() < sense.sensmap.dim()


class UnrolledCGSENSE(nn.Module):
    """CG-SENSE with ``niter`` unrolled iterations as a differentiable module.

    forward(kspace): kspace (ncoils, [ncontrasts], nviews, nsamples) ->
    image (*N) (or (ncontrasts, *N)).
    """

    def __init__(self, coord, shape, sensmap, niter=10, lamda=0.0, dcf=None, device=None, width=4,
                 oversamp=1.25):
        super().__init__()
        self.E, self.EHE = encoding_op(coord, shape, sensmap, dcf, device, width, oversamp)
        self.niter = niter
        self.lamda = lamda
        self.dcf = dcf

    def forward(self, kspace):
        return cg_sense(kspace, self.E, self.EHE, self.niter, self.lamda, self.dcf)


def capture_unrolled(model, sample_kspace, warmup=3):
    """Capture ``model``'s forward and backward (an UnrolledCGSENSE, or any
    module over the same operators) into HIP graphs with
    torch.cuda.make_graphed_callables and return it: the module's forward
    is replaced in place by the graphed one, so later calls replay the
    captured kernels instead of launching them one by one from Python.

    ``sample_kspace`` fixes what later calls must match: shape, dtype,
    device and requires_grad. The native launches take no host
    synchronisation and record no memsets (plans and launch tables are built
    before capture), so the captured work is the same kernels on the same
    buffers. Capture the model on the stream it will run on; do not rebuild
    its plans afterwards."""
    if not isinstance(sample_kspace, torch.Tensor) or sample_kspace.device.type != "cuda":
        raise ValueError("sample_kspace must be a HIP device tensor")
    sample = sample_kspace.detach().clone().requires_grad_(sample_kspace.requires_grad)
    return torch.cuda.make_graphed_callables(model, (sample,), num_warmup_iters=warmup)
