"""CG solver and unrolled CG steps (CPU only).

Restates deepmr's tests/optim/test_optim.py:40-55 (CG against the closed-form
Tikhonov solution) and checks that unrolled CGStep iterations are
differentiable (float64 gradcheck on a dense SPD operator).
"""

import itertools

import numpy as np
import pytest
import torch

from tilenufft import linops, optim


def _setup(n, lamda, dtype):
    A = (torch.eye(n) + 0.1 * torch.ones(n, n)).to(dtype)
    x = torch.arange(n).to(dtype)
    y = A @ x
    x_ref = torch.linalg.solve(A.T @ A + lamda * torch.eye(n, dtype=dtype), A.T @ y)
    return A, x_ref, y


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64, torch.complex64, torch.complex128])
def test_cg_matches_closed_form(dtype):
    n, lamda = 5, 0.1
    A, x_ref, y = _setup(n, lamda, dtype)
    x, _ = optim.cg_solve(A.T @ y, lambda v: A.T @ A @ v, niter=1000, lamda=lamda, ndim=1)
    np.testing.assert_allclose(x.numpy(), x_ref.numpy(), rtol=1e-4, atol=1e-4)


def test_cg_with_linop_and_history():
    n = 6
    M = torch.randn(n, n, dtype=torch.float64, generator=torch.Generator().manual_seed(0))
    M = M @ M.T + n * torch.eye(n, dtype=torch.float64)

    class Dense(linops.Linop):
        def forward(self, v):
            return M @ v

        def _adjoint_linop(self):
            return self

    b = torch.arange(n, dtype=torch.float64)
    x, hist = optim.cg_solve(b, Dense(1), niter=50, tol=0.0, save_history=True)
    np.testing.assert_allclose(x.numpy(), torch.linalg.solve(M, b).numpy(), rtol=1e-8)
    assert len(hist) > 0


def test_unrolled_cgstep_gradcheck():
    n = 4
    g = torch.Generator().manual_seed(1)
    M = torch.randn(n, n, dtype=torch.complex128, generator=g)
    M = M @ M.conj().T + n * torch.eye(n, dtype=torch.complex128)

    def unrolled(b):
        step = optim.CGStep(lambda v: M @ v, b, ndim=1)
        x = torch.zeros_like(b)
        for _ in range(3):
            x = step(x)
        return x

    b = torch.randn(n, dtype=torch.complex128, generator=g, requires_grad=True)
    assert torch.autograd.gradcheck(unrolled, (b,))


def test_linop_algebra():
    class Scale(linops.Linop):
        def __init__(self, a):
            super().__init__(1)
            self.a = a

        def forward(self, v):
            return self.a * v

        def _adjoint_linop(self):
            return Scale(np.conj(self.a))

    A, B = Scale(2.0 + 1.0j), Scale(3.0)
    v = torch.ones(3, dtype=torch.complex128)
    assert torch.allclose((A * B)(v), (2 + 1j) * 3 * v)
    assert torch.allclose((A + B)(v), (5 + 1j) * v)
    assert torch.allclose((A - B)(v), (-1 + 1j) * v)
    assert torch.allclose((A * B).H(v), (2 - 1j) * 3 * v)
    assert torch.allclose((0.5 * A)(v), (1 + 0.5j) * v)
    assert torch.allclose(linops.Identity(1)(v), v)


@pytest.mark.parametrize("ndim, multicontrast", list(itertools.product([2, 3], [True, False])))
def test_sense_ops_are_adjoint(ndim, multicontrast):
    g = torch.Generator().manual_seed(2)
    shape = (6,) * ndim
    ncoil = 3
    smap = torch.randn((1, ncoil, *shape), dtype=torch.complex128, generator=g)
    S = linops.SenseOp(ndim, smap, multicontrast=multicontrast)
    x_shape = (2, *shape) if multicontrast else shape
    x = torch.randn(x_shape, dtype=torch.complex128, generator=g)
    y = S(x)
    r = torch.randn(y.shape, dtype=torch.complex128, generator=g)
    lhs = torch.vdot(y.flatten(), r.flatten())
    rhs = torch.vdot(x.flatten(), S.H(r).flatten())
    assert torch.allclose(lhs, rhs)
