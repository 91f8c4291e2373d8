"""Native CG iteration (tn_cg_iterate / tn_cg_iterate_backward) behind
CGStep (src/deepmr/optim/cg.py:150-221), GPU.

The native path runs for complex64 systems on a HIP device; the torch
expressions of CGStep (the path for other dtypes) are the reference. On a
random Hermitian positive-definite operator (batched matrix-vector products,
so autograd carries q = A p's gradient through A), several iterations from
zero are compared with the same recursion in complex128: iterates, residual
norms and the gradients of a loss of the iterate with respect to the
right-hand side and to the operator's matrix. Leading axes are independent
systems; 1D and 2D system shapes.
"""

import math

import pytest
import torch

from tilenufft.optim import CGStep

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def cplx(shape, seed, dtype=torch.complex64):
    g = torch.Generator().manual_seed(seed)
    return torch.complex(torch.randn(shape, generator=g), torch.randn(shape, generator=g)).to(dtype)


def rel(a, b):
    a, b = a.detach().to(torch.complex128), b.detach().to(torch.complex128)
    return float((a - b).abs().norm() / b.abs().norm())


def run(y, M, ndim, niter, dtype):
    n = M.shape[-1]

    def AHA(v):
        return (v.reshape(-1, n) @ M.transpose(0, 1).contiguous()).reshape(v.shape)

    step = CGStep(AHA, y, ndim=ndim)
    x = torch.zeros_like(y)
    for _ in range(niter):
        x = step(x)
    return x, step


@pytest.mark.parametrize("lead, img", [((3,), (300,)), ((2,), (24, 20))])
def test_native_cg_matches_complex128_recursion(lead, img):
    ndim = len(img)
    n = math.prod(img)
    print(f"native CG lead={lead} img={img}", flush=True)
    Q = cplx((n, n), 1) / math.sqrt(n)
    M0 = (Q @ Q.conj().T + 0.5 * torch.eye(n)).to(torch.complex128)  # Hermitian, cond ~ 10
    y0 = cplx((*lead, *img), 2, torch.complex128)
    target = cplx((*lead, *img), 3, torch.complex128)
    res = {}
    for dt in (torch.complex64, torch.complex128):
        M = M0.to(dt).to(DEV).requires_grad_(True)
        y = y0.to(dt).to(DEV).requires_grad_(True)
        x, step = run(y, M, ndim, 6, dt)
        assert step._native == (dt == torch.complex64)
        loss = (x - target.to(dt).to(DEV)).abs().pow(2).sum() + step.rsnew.sum()
        loss.backward()
        res[dt] = (x, step.rsnew, y.grad, M.grad)
    x32, rs32, gy32, gM32 = res[torch.complex64]
    x64, rs64, gy64, gM64 = res[torch.complex128]
    assert rs32.shape == rs64.shape == (*lead, *([1] * ndim))
    print(f"  x {rel(x32, x64):.1e} rs {rel(rs32, rs64):.1e} grad y {rel(gy32, gy64):.1e} grad M {rel(gM32, gM64):.1e}",
          flush=True)
    assert rel(x32, x64) < 1e-4
    assert rel(rs32, rs64) < 1e-4
    assert rel(gy32, gy64) < 1e-3
    assert rel(gM32, gM64) < 1e-3


def test_native_cg_equals_torch_path_in_float32():
    """The same complex64 system through the native iteration and through
    CGStep's torch expressions (native switched off): equal to float32
    round-off, values and gradients."""
    print("native CG vs torch CG, complex64", flush=True)
    n = 200
    Q = cplx((n, n), 4) / math.sqrt(n)
    M = (Q @ Q.conj().T + 0.5 * torch.eye(n)).to(DEV)
    out = []
    for native in (True, False):
        y = cplx((4, n), 5).to(DEV).requires_grad_(True)

        def AHA(v):
            return v @ M.transpose(0, 1).contiguous()

        step = CGStep(AHA, y, ndim=1)
        step._native = native
        x = torch.zeros_like(y)
        for _ in range(5):
            x = step(x)
        x.abs().pow(2).sum().backward()
        out.append((x.detach(), y.grad))
    assert rel(out[0][0], out[1][0]) < 1e-5
    assert rel(out[0][1], out[1][1]) < 1e-4
