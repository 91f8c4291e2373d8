"""GPU parity of the register-tile gridding (grid_mode 4, 3D).

One wave per 8x8x4 gridding tile sums its cells in registers over the
tile's samples in ascending order, with the same per-sample taps and the
same (wz wy) wx weight as the sparse-matrix gridding. Checked against the
oracle (relative L2 <= 1e-5 for complex64), bitwise against the matrix
gridding where neither splits its work (both then add each cell's terms in
the same order with the same operations), bitwise run to run, with split
tiles, coil chunks, footprints that wrap around the grid edge, widths 2-4,
an empty trajectory, and the low-rank adjoint, which takes the LDS-atomic
kernel over the same bins.
"""

import numpy as np
import pytest
import torch

import tilenufft
from oracle import nufft_oracle as orc
from tilenufft import testing

pytestmark = pytest.mark.gpu

RTOL = 1e-5
DEV = "cuda:0"


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def cnormal(shape, seed):
    return testing.complex_normal(shape, seed=seed).numpy()


def uniform3d(nspokes, nsamp, n, seed):
    """Uniform samples over the whole k-space cube: footprints at the faces
    wrap around the grid edge."""
    rng = np.random.default_rng(seed)
    return rng.uniform(-n / 2, n / 2, (nspokes, nsamp, 3)).astype(np.float32)


CASES = [
    # (N, width, osf, coils, grid_split, trajectory)
    (16, 4, 2.0, 3, 0, "koosh"),      # 32^3 grid, dense centre
    (16, 4, 2.0, 8, 64, "koosh"),     # every hot tile split into parts
    (20, 4, 2.0, 12, 0, "koosh"),     # 40^3: 5 tiles per x / y axis; chunks of 8 + 4 coils
    (24, 3, 2.0, 2, 16, "uniform"),   # width 3, tiny parts, wrapped footprints
    (16, 2, 2.0, 1, 0, "uniform"),    # width 2, one coil (2-coil records)
    (32, 4, 1.5, 5, 0, "uniform"),    # 48^3
]


@pytest.mark.parametrize("n, width, osf, ncoil, split, traj", CASES)
def test_tiles_gridding_matches_oracle_and_is_deterministic(n, width, osf, ncoil, split, traj):
    print(f"tiles N={n} W={width} osf={osf} coils={ncoil} split={split} {traj}", flush=True)
    coord = testing.kooshball(60, 32, n) if traj == "koosh" else uniform3d(40, 50, n, 7)
    k = cnormal((ncoil,) + coord.shape[:-1], 70 + width)
    w = np.random.default_rng(71).uniform(0.5, 1.5, coord.shape[:-1]).astype(np.float32)
    plan = tilenufft.fft.plan_nufft(coord, n, width, osf, device=DEV, grid_mode=4, grid_split=split)
    assert plan.interpolator.grid_mode_effective() == 4
    kt = torch.from_numpy(k).to(DEV)
    wt = torch.from_numpy(w)
    a = tilenufft.fft.apply_nufft_adj(kt, plan, weight=wt)
    b = tilenufft.fft.apply_nufft_adj(kt, plan, weight=wt)
    torch.cuda.synchronize()
    assert torch.equal(a.view(torch.float32), b.view(torch.float32))
    ref = orc.apply_nufft_adj(k, orc.plan_nufft(coord, n, width, osf), weight=w)
    assert rel(a.cpu().numpy(), ref) < RTOL


@pytest.mark.parametrize("chunk", [1, 3, 8])
def test_tiles_coil_chunks(chunk):
    """1, 3 and 8 coils per item over 11 coils (records of 2, 4 and 8 coils;
    the last chunk partial)."""
    n = 16
    print(f"tiles chunk={chunk}", flush=True)
    coord = testing.kooshball(50, 32, n)
    k = cnormal((11,) + coord.shape[:-1], 72)
    plan = tilenufft.fft.plan_nufft(coord, n, 4, 2.0, device=DEV, grid_mode=4, grid_chunk=chunk)
    a = tilenufft.fft.apply_nufft_adj(torch.from_numpy(k).to(DEV), plan).cpu().numpy()
    ref = orc.apply_nufft_adj(k, orc.plan_nufft(coord, n, 4, 2.0))
    assert rel(a, ref) < RTOL


def test_tiles_bitwise_equal_to_matrix_without_splits():
    """With no split tiles and no long matrix rows, both variants add each
    cell's terms (wz wy) wx * scaled sample in ascending sample order: the
    adjoints are the same bits, for 8- and 3-coil records."""
    n = 16
    coord = uniform3d(30, 100, n, 9)
    for ncoil in (8, 3):
        print(f"tiles vs matrix coils={ncoil}", flush=True)
        k = torch.from_numpy(cnormal((ncoil,) + coord.shape[:-1], 73)).to(DEV)
        pt = tilenufft.fft.plan_nufft(coord, n, 4, 2.0, device=DEV, grid_mode=4, grid_split=-1)
        pm = tilenufft.fft.plan_nufft(coord, n, 4, 2.0, device=DEV, grid_mode=3, grid_split=-1)
        a = tilenufft.fft.apply_nufft_adj(k, pt)
        b = tilenufft.fft.apply_nufft_adj(k, pm)
        torch.cuda.synchronize()
        assert torch.equal(a.view(torch.float32), b.view(torch.float32))


def test_tiles_forward_unchanged_and_dot_product():
    """The forward transform of a tiles plan (interpolation is unchanged)
    against the oracle, and the adjoint pair <A x, k> = prod(os) <x, A^H k>."""
    n = 16
    coord = testing.kooshball(40, 32, n)
    x = cnormal((2, n, n, n), 74)
    plan = tilenufft.fft.plan_nufft(coord, n, 4, 2.0, device=DEV, grid_mode=4)
    y = tilenufft.fft.apply_nufft(torch.from_numpy(x).to(DEV), plan).cpu().numpy()
    p_ref = orc.plan_nufft(coord, n, 4, 2.0)
    assert rel(y, orc.apply_nufft(x, p_ref)) < RTOL
    k = cnormal(y.shape, 75)
    xa = tilenufft.fft.apply_nufft_adj(torch.from_numpy(k).to(DEV), plan).cpu().numpy()
    lhs = np.vdot(y.astype(np.complex128), k.astype(np.complex128))
    rhs = float(np.prod(plan.os_shape)) * np.vdot(x.astype(np.complex128), xa.astype(np.complex128))
    assert abs(lhs - rhs) / (np.linalg.norm(y) * np.linalg.norm(k)) < 1e-6


def test_tiles_empty_trajectory():
    print("tiles empty trajectory", flush=True)
    coord = np.zeros((0, 8, 3), np.float32)
    plan = tilenufft.fft.plan_nufft(coord, 16, 4, 2.0, device=DEV, grid_mode=4)
    img = tilenufft.fft.apply_nufft_adj(torch.zeros((2, 0, 8), dtype=torch.complex64, device=DEV), plan)
    torch.cuda.synchronize()
    assert img.shape == (2, 16, 16, 16) and float(img.abs().max()) == 0.0


def test_tiles_refused_where_they_do_not_fit():
    """2D plans, widths above 4 and grids that are not whole tiles are
    refused with a message; the automatic mode never picks them there."""
    with pytest.raises(RuntimeError, match="register tiles"):
        tilenufft.fft.plan_nufft(testing.vd_spiral(4, 64, 32), 32, 4, 2.0, device=DEV, grid_mode=4)
    with pytest.raises(RuntimeError, match="register tiles"):
        tilenufft.fft.plan_nufft(testing.kooshball(8, 16, 16), 16, 6, 2.0, device=DEV, grid_mode=4)
    with pytest.raises(RuntimeError, match="register tiles"):
        tilenufft.fft.plan_nufft(testing.kooshball(8, 16, 14), 14, 4, 1.5, device=DEV, grid_mode=4)  # 21^3


def test_autotune_3d_considers_register_tiles():
    """The 3D autotuner's gridding candidates include the register tiles
    (tile-independent, with their own split candidates); it leaves the plan
    on the fastest variant, whose adjoint matches the oracle."""
    from tilenufft.autotune import default_candidates, tune_batch
    from tilenufft.batch import BatchedNUFFT

    assert 4 in default_candidates(3)["grid_modes"] and 4 not in default_candidates(2)["grid_modes"]

    n = 16
    coord = testing.kooshball(40, 32, n)
    plan = tilenufft.fft.plan_nufft(coord, n, 4, 2.0, device=DEV)
    best = tune_batch([plan], 3, reps=2, use_cache=False)
    assert best["grid_mode"] in (1, 2, 3, 4)
    assert plan.interpolator.grid_mode_effective() == best["grid_mode"]
    k = cnormal((3,) + coord.shape[:-1], 76)
    a = tilenufft.fft.apply_nufft_adj(torch.from_numpy(k).to(DEV), plan).cpu().numpy()
    assert rel(a, orc.apply_nufft_adj(k, orc.plan_nufft(coord, n, 4, 2.0))) < RTOL
    b = BatchedNUFFT([plan], 3)
    assert b.grid.shape[-3:] == (2 * n,) * 3


def test_tiles_batched_plans_and_frames_equal_single_plans():
    """Two register-tile plans (one of two frames, one of one) with 5 and 8
    coils in one batched gridding launch: each plan's grid is the same bits
    as its own launch (per-job taps, bins and coil-minor copies), and the
    two-frame plan's adjoint matches the oracle frame by frame."""
    from tilenufft.batch import BatchedNUFFT

    n = 16
    c1 = np.stack([testing.kooshball(30, 32, n), uniform3d(30, 32, n, 11)])  # (2 frames, views, samples, 3)
    c2 = testing.kooshball(45, 32, n)
    p1 = tilenufft.fft.plan_nufft(c1, n, 4, 2.0, device=DEV, grid_mode=4, grid_split=48)
    p2 = tilenufft.fft.plan_nufft(c2, n, 4, 2.0, device=DEV, grid_mode=4)
    assert p1.nframes == 2
    b = BatchedNUFFT([p1, p2], [5, 8])
    gen = torch.Generator().manual_seed(77)
    b.kspace.copy_(testing.complex_normal(tuple(b.kspace.shape), generator=gen).to(DEV))
    b.gridding(b.grid)
    torch.cuda.synchronize()
    for k, (p, c) in enumerate([(p1, 5), (p2, 8)]):
        print(f"tiles batch plan {k}", flush=True)
        single = BatchedNUFFT([p], c)
        single.kspace.copy_(b.kspace_view(k).reshape(-1))
        single.gridding(single.grid)
        torch.cuda.synchronize()
        g0 = b.grid_off[k] // int(np.prod(b.os_shape))
        mine = b.grid[g0:g0 + c * p.nframes]
        assert torch.equal(mine.view(torch.float32), single.grid.view(torch.float32))
    ks = b.kspace_view(0).cpu().numpy()  # (coils, frames, npts)
    img = tilenufft.fft.apply_nufft_adj(torch.from_numpy(ks.reshape(5, 2, 30, 32)).to(DEV), p1).cpu().numpy()
    ref = orc.apply_nufft_adj(ks.reshape(5, 2, 30, 32), orc.plan_nufft(c1, n, 4, 2.0))
    assert rel(img, ref) < RTOL
