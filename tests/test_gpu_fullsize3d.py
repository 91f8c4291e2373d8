"""Full-size 3D kooshball (GPU): 128^3 image, 8 coils, 20000 spokes x 256
samples, KB width 4, osf 2 (256^3 grid, the automatic gridding matrix of
327 M entries). Too large for the oracle, so size-independent properties:

- the dot-product identity <A x, y> = prod(os_shape) <x, nufft_adj(y)>
  (deepmr's adjoint is A^H / prod(os_shape)), in float64;
- bitwise determinism of the forward and adjoint on repeated calls;
- linearity of the adjoint: nufft_adj(a y1 + y2) = a nufft_adj(y1) + nufft_adj(y2).
"""

import numpy as np
import pytest
import torch

import tilenufft
from tilenufft import testing

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def test_kooshball_full_size_properties():
    print("full-size 3D kooshball 128^3, 8 coils, 20000 x 256, osf 2", flush=True)
    coord = testing.kooshball(20000, 256, 128)
    plan = tilenufft.fft.plan_nufft(coord, 128, 4, 2.0, device=DEV)
    assert plan.os_shape == (256, 256, 256) and plan.grid_mode in (0, 3)
    gen = torch.Generator().manual_seed(3)
    x = testing.complex_normal((8, 128, 128, 128), generator=gen).to(DEV)
    k = testing.complex_normal((8, 20000, 256), generator=gen).to(DEV)
    y = tilenufft.fft.apply_nufft(x, plan)
    xa = tilenufft.fft.apply_nufft_adj(k, plan)
    torch.cuda.synchronize()
    lhs = torch.vdot(y.reshape(-1).to(torch.complex128), k.reshape(-1).to(torch.complex128))
    rhs = float(np.prod(plan.os_shape)) * torch.vdot(x.reshape(-1).to(torch.complex128),
                                                     xa.reshape(-1).to(torch.complex128))
    err = float(abs(lhs - rhs) / abs(lhs))
    print(f"  dot identity rel err {err:.2e}", flush=True)
    assert err < 1e-4
    y2 = tilenufft.fft.apply_nufft(x, plan)
    xa2 = tilenufft.fft.apply_nufft_adj(k, plan)
    torch.cuda.synchronize()
    assert torch.equal(y.view(torch.float32), y2.view(torch.float32))
    assert torch.equal(xa.view(torch.float32), xa2.view(torch.float32))
    del y2, xa2
    k2 = testing.complex_normal((8, 20000, 256), generator=gen).to(DEV)
    a = 0.37 - 1.1j
    lin = tilenufft.fft.apply_nufft_adj(a * k + k2, plan)
    ref = a * xa + tilenufft.fft.apply_nufft_adj(k2, plan)
    e = float(torch.linalg.vector_norm(lin - ref) / torch.linalg.vector_norm(ref))
    print(f"  adjoint linearity rel err {e:.2e}", flush=True)
    assert e < 1e-5


def test_kooshball_beyond_matrix_budget_uses_register_tiles():
    """36,000 spokes x 256 samples: the gridding matrix would hold 590 M
    entries (4.7 GB, above the automatic mode's 4 GiB), so the automatic plan
    grids through the register tiles; the full-size adjoint then agrees with
    an explicit matrix plan of the same trajectory to FP32 rounding (the
    same per-cell order where neither splits; different partitions of the
    hot cells), and satisfies the dot-product identity."""
    print("kooshball 36000 x 256 beyond the matrix budget", flush=True)
    coord = testing.kooshball(36000, 256, 128)
    plan = tilenufft.fft.plan_nufft(coord, 128, 4, 2.0, device=DEV)
    assert plan.interpolator.grid_mode_effective() == 4
    assert plan.interpolator.device_bytes() < 1.5e9
    gen = torch.Generator().manual_seed(5)
    k = testing.complex_normal((2, 36000, 256), generator=gen).to(DEV)
    x = testing.complex_normal((2, 128, 128, 128), generator=gen).to(DEV)
    xa = tilenufft.fft.apply_nufft_adj(k, plan)
    y = tilenufft.fft.apply_nufft(x, plan)
    torch.cuda.synchronize()
    lhs = torch.vdot(y.reshape(-1).to(torch.complex128), k.reshape(-1).to(torch.complex128))
    rhs = float(np.prod(plan.os_shape)) * torch.vdot(x.reshape(-1).to(torch.complex128),
                                                     xa.reshape(-1).to(torch.complex128))
    err = float(abs(lhs - rhs) / abs(lhs))
    print(f"  dot identity rel err {err:.2e}", flush=True)
    assert err < 1e-4
    del plan, y
    torch.cuda.empty_cache()
    pm = tilenufft.fft.plan_nufft(coord, 128, 4, 2.0, device=DEV, grid_mode=3)
    xm = tilenufft.fft.apply_nufft_adj(k, pm)
    torch.cuda.synchronize()
    e = float(torch.linalg.vector_norm(xa - xm) / torch.linalg.vector_norm(xm))
    print(f"  register tiles vs matrix rel err {e:.2e}", flush=True)
    assert e < 1e-5
