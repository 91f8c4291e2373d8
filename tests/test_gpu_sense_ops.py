"""SENSE coil operators on the GPU (native coil combination).

SenseOp / SenseAdjointOp (src/deepmr/linops/coil.py:11-205) take a native
path for constant maps on a HIP device: the expansion is one broadcast
product and the combination (and the expansion's backward) one native
kernel, tn_coil_combine. Checked against deepmr's torch expressions,
evaluated in complex128, for values and for the gradients of a loss, over
the layouts the operators accept: one or two map sets, a slice axis,
multicontrast images, 2D and 3D; real input; and maps that need gradients
(the torch path).
"""

import numpy as np
import pytest
import torch

from tilenufft import linops

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def cplx(shape, seed, dtype=torch.complex64):
    g = torch.Generator().manual_seed(seed)
    return torch.complex(torch.randn(shape, generator=g), torch.randn(shape, generator=g)).to(dtype)


def rel(a, b):
    return float((a.to(torch.complex128) - b.to(torch.complex128)).abs().norm() / b.abs().norm())


CASES = [  # ndim, smap shape, image shape, multicontrast
    (2, (1, 8, 32, 32), (32, 32), False),
    (2, (2, 3, 4, 24, 20), (3, 24, 20), False),
    (2, (1, 3, 4, 24, 20), (3, 2, 24, 20), True),
    (3, (2, 5, 8, 10, 12), (8, 10, 12), False),
]


@pytest.mark.parametrize("ndim, smap_shape, img_shape, mc", CASES)
def test_sense_ops_match_torch_expressions(ndim, smap_shape, img_shape, mc):
    print(f"SENSE ops ndim={ndim} maps={smap_shape} image={img_shape} mc={mc}", flush=True)
    smap = cplx(smap_shape, 1).to(DEV)
    x = cplx(img_shape, 2).to(DEV).requires_grad_(True)
    S = linops.SenseOp(ndim, smap, multicontrast=mc)
    y = S(x)
    # deepmr's expressions in complex128
    s64 = smap.to(torch.complex128)
    if mc:
        s64 = s64.unsqueeze(-ndim - 1)
    x64 = x.detach().to(torch.complex128).requires_grad_(True)
    y64 = s64 * x64.unsqueeze(-ndim - 2 if mc else -ndim - 1)
    assert y.shape == y64.shape and rel(y.detach(), y64.detach()) < 1e-6
    w = cplx(tuple(y.shape), 3).to(DEV)
    loss = (w.conj() * y).real.sum() + y.abs().pow(2).sum()
    loss.backward()
    loss64 = (w.to(torch.complex128).conj() * y64).real.sum() + y64.abs().pow(2).sum()
    loss64.backward()
    assert rel(x.grad, x64.grad) < 1e-5
    # adjoint: combination of coil images
    yy = cplx(tuple(y.shape), 4).to(DEV).requires_grad_(True)
    xa = S.H(yy)
    yy64 = yy.detach().to(torch.complex128).requires_grad_(True)
    coil = -ndim - 2 if mc else -ndim - 1
    xa64 = (s64.conj() * yy64).sum(dim=coil).sum(dim=0)
    assert xa.shape == xa64.shape and rel(xa.detach(), xa64.detach()) < 1e-6
    v = cplx(tuple(xa.shape), 5).to(DEV)
    (v.conj() * xa).real.sum().backward()
    (v.to(torch.complex128).conj() * xa64).real.sum().backward()
    assert rel(yy.grad, yy64.grad) < 1e-6


def test_sense_ops_real_input_and_maps_with_gradients():
    print("SENSE ops: real input, maps needing gradients", flush=True)
    smap = cplx((1, 4, 16, 16), 6).to(DEV)
    x = torch.randn(16, 16, device=DEV)
    y = linops.SenseOp(2, smap, multicontrast=False)(x)
    assert rel(y, smap * x) < 1e-6
    sg = smap.clone().requires_grad_(True)  # the torch expression, gradients for the maps too
    yy = cplx((1, 4, 16, 16), 7).to(DEV)
    xa = linops.SenseAdjointOp(2, sg, multicontrast=False)(yy)
    assert rel(xa.detach(), (smap.conj() * yy).sum(1).sum(0)) < 1e-6
    xa.abs().pow(2).sum().backward()
    assert sg.grad is not None and torch.isfinite(sg.grad.abs()).all()
    # numpy in, numpy out
    out = linops.SenseOp(2, smap, multicontrast=False)(np.ones((16, 16), np.complex64))
    assert isinstance(out, np.ndarray) and out.shape == (1, 4, 16, 16)
