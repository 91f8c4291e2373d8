"""GPU parity of the NUFFT path against the oracle (deepmr restatement).

Bar: relative L2 <= 1e-5 for complex64 against the oracle on the same seeded
inputs; bitwise run-to-run determinism of the adjoint; dot-product test with
deepmr's 1/prod(os_shape) adjoint scaling; error vs the FP64 NDFT within 1%
of the oracle's own error.
"""

import numpy as np
import pytest
import torch

import tilenufft
from oracle import ndft
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


def run_pair(x, coord, shape, width, osf, weight=None):
    plan = tilenufft.fft.plan_nufft(coord, shape, width, osf, device=DEV)
    y = tilenufft.fft.apply_nufft(torch.from_numpy(x).to(DEV), plan, weight=weight)
    k = cnormal(y.shape, 99)
    xa = tilenufft.fft.apply_nufft_adj(torch.from_numpy(k).to(DEV), plan, weight=weight)
    return plan, y.cpu().numpy(), k, xa.cpu().numpy()


CASES_2D = [
    # (N, coils, nviews, nsamples, osf)
    (64, 4, 32, 128, 2.0),
    (64, 3, 24, 96, 1.25),
    (48, 2, 20, 100, 1.5),
]


@pytest.mark.parametrize("n, ncoil, nviews, nsamp, osf", CASES_2D)
def test_nufft2d_matches_oracle(n, ncoil, nviews, nsamp, osf):
    print(f"2D radial N={n} coils={ncoil} {nviews}x{nsamp} osf={osf}", flush=True)
    coord = testing.golden_angle_radial(nviews, nsamp, n)
    x = cnormal((ncoil, n, n), 1)
    plan, y, k, xa = run_pair(x, coord, n, 4, osf)
    p_ref = orc.plan_nufft(coord, n, 4, osf)
    y_ref = orc.apply_nufft(x, p_ref)
    xa_ref = orc.apply_nufft_adj(k, p_ref)
    assert y.shape == y_ref.shape and xa.shape == xa_ref.shape
    assert rel(y, y_ref) < RTOL
    assert rel(xa, xa_ref) < RTOL


def test_nufft2d_anisotropic_odd():
    # odd and unequal axes catch x/y swaps and centring off-by-ones
    shape = (30, 45)  # (ny, nx)
    print("2D anisotropic shape=(30,45) osf=1.5", flush=True)
    rng = np.random.default_rng(3)
    coord = np.stack([rng.uniform(-22.5, 22.5, (7, 50)), rng.uniform(-15, 15, (7, 50))], -1).astype(np.float32)
    x = cnormal((2, *shape), 2)
    plan, y, k, xa = run_pair(x, coord, shape, 4, 1.5)
    p_ref = orc.plan_nufft(coord, shape, 4, 1.5)
    assert plan.os_shape == (45, 68)
    assert rel(y, orc.apply_nufft(x, p_ref)) < RTOL
    assert rel(xa, orc.apply_nufft_adj(k, p_ref)) < RTOL


@pytest.mark.parametrize("osf", [1.5, 2.0])
def test_nufft3d_matches_oracle(osf):
    n = 16
    print(f"3D kooshball N={n} osf={osf}", flush=True)
    coord = testing.kooshball(60, 32, n)
    x = cnormal((2, n, n, n), 4)
    plan, y, k, xa = run_pair(x, coord, n, 4, osf)
    p_ref = orc.plan_nufft(coord, n, 4, osf)
    assert rel(y, orc.apply_nufft(x, p_ref)) < RTOL
    assert rel(xa, orc.apply_nufft_adj(k, p_ref)) < RTOL


@pytest.mark.parametrize("ndim, n, tile", [(2, 20, (14, 14)), (2, 20, (5, 3)), (3, 12, None),
                                           (3, 12, (7, 5, 16))])
def test_narrow_edge_tiles(ndim, n, tile):
    # grids whose last tile is narrower than the kernel: footprints wrap back
    # into a tile they already touched (regression: such samples were binned
    # twice into that tile)
    print(f"narrow edge tiles ndim={ndim} N={n} tile={tile}", flush=True)
    coord = testing.golden_angle_radial(16, 40, n) if ndim == 2 else testing.kooshball(40, 24, n)
    x = cnormal((2,) + (n,) * ndim, 21)
    plan = tilenufft.fft.plan_nufft(coord, n, 4, 1.5, device=DEV, tile=tile)
    y = tilenufft.fft.apply_nufft(torch.from_numpy(x).to(DEV), plan).cpu().numpy()
    k = cnormal(y.shape, 22)
    xa = tilenufft.fft.apply_nufft_adj(torch.from_numpy(k).to(DEV), plan).cpu().numpy()
    p_ref = orc.plan_nufft(coord, n, 4, 1.5)
    assert rel(y, orc.apply_nufft(x, p_ref)) < RTOL
    assert rel(xa, orc.apply_nufft_adj(k, p_ref)) < RTOL


def test_multiframe_and_weight():
    # 3 frames (contrasts) with their own trajectories, batch (2 slices, 3 coils)
    n, nf = 32, 3
    print("2D multiframe nf=3 batch=(2,3) with weight", flush=True)
    coord = np.stack([testing.golden_angle_radial(10, 64, n, offset=0.3 * f) for f in range(nf)])
    x = cnormal((2, 3, nf, n, n), 5)
    w = np.random.default_rng(6).uniform(0.5, 1.5, coord.shape[:-1]).astype(np.float32)
    plan, y, k, xa = run_pair(x, coord, n, 4, 2.0, weight=torch.from_numpy(w))
    p_ref = orc.plan_nufft(coord, n, 4, 2.0)
    assert y.shape == (2, 3, nf, 10, 64)
    assert rel(y, orc.apply_nufft(x, p_ref, weight=w)) < RTOL
    assert rel(xa, orc.apply_nufft_adj(k, p_ref, weight=w)) < RTOL


def test_multiframe_3d_with_weight():
    """3D with 2 frames of their own kooshball trajectories (one rotated),
    2 coils and a density weight: each frame's interpolation entries are
    sorted by (frame, tile, footprint origin), so the frames' samples never
    mix; forward and adjoint match the oracle."""
    n, nf = 16, 2
    print("3D multiframe nf=2 coils=2 with weight", flush=True)
    base = testing.kooshball(60, 24, n)
    coord = np.stack([base, base[..., [1, 2, 0]]])  # (frames, views, samples, 3)
    x = cnormal((2, nf, n, n, n), 7)
    w = np.random.default_rng(8).uniform(0.5, 1.5, coord.shape[:-1]).astype(np.float32)
    plan, y, k, xa = run_pair(x, coord, n, 4, 2.0, weight=torch.from_numpy(w))
    p_ref = orc.plan_nufft(coord, n, 4, 2.0)
    assert y.shape == (2, nf, 60, 24)
    assert rel(y, orc.apply_nufft(x, p_ref, weight=w)) < RTOL
    assert rel(xa, orc.apply_nufft_adj(k, p_ref, weight=w)) < RTOL


def test_dot_product_2d_and_3d():
    for ndim, n in ((2, 40), (3, 12)):
        print(f"dot test ndim={ndim} N={n}", flush=True)
        coord = testing.golden_angle_radial(12, 80, n) if ndim == 2 else testing.kooshball(40, 24, n)
        x = cnormal((2,) + (n,) * ndim, 7)
        plan, y, k, xa = run_pair(x, coord, n, 4, 1.5)
        lhs = np.vdot(y.astype(np.complex128), k.astype(np.complex128))
        rhs = np.prod(plan.os_shape) * np.vdot(x.astype(np.complex128), xa.astype(np.complex128))
        assert abs(lhs - rhs) / abs(lhs) < 1e-5


def test_adjoint_bitwise_deterministic():
    n = 64
    print("determinism 2D N=64 8 coils dense radial", flush=True)
    coord = testing.golden_angle_radial(64, 256, n)
    plan = tilenufft.fft.plan_nufft(coord, n, 4, 2.0, device=DEV)
    k = torch.from_numpy(cnormal((8, 64, 256), 8)).to(DEV)
    a = tilenufft.fft.apply_nufft_adj(k, plan)
    b = tilenufft.fft.apply_nufft_adj(k, plan)
    torch.cuda.synchronize()
    assert torch.equal(a.view(torch.float32), b.view(torch.float32))


@pytest.mark.parametrize("ndim", [2, 3])
def test_ndft_error_matches_oracle(ndim):
    n = 24 if ndim == 2 else 10
    print(f"NDFT error ratio ndim={ndim}", flush=True)
    coord = testing.golden_angle_radial(20, 48, n) if ndim == 2 else testing.kooshball(30, 20, n)
    x = cnormal((n,) * ndim, 9)
    plan = tilenufft.fft.plan_nufft(coord, n, 4, 2.0, device=DEV)
    y = tilenufft.fft.apply_nufft(torch.from_numpy(x).to(DEV), plan).cpu().numpy()
    y_ref = orc.nufft(x, coord, n, 4, 2.0)
    exact = ndft.ndft(x, coord, (n,) * ndim).reshape(y.shape)
    e_gpu, e_ref = rel(y, exact), rel(y_ref, exact)
    assert abs(e_gpu / e_ref - 1) < 0.01
    # adjoint
    k = cnormal(y.shape, 10)
    xa = tilenufft.fft.apply_nufft_adj(torch.from_numpy(k).to(DEV), plan).cpu().numpy()
    xa_ref = orc.nufft_adj(k, coord, n, 4, 2.0)
    exact_a = ndft.ndft_adjoint(k.reshape(-1), coord, (n,) * ndim) / np.prod(plan.os_shape)
    assert abs(rel(xa, exact_a) / rel(xa_ref, exact_a) - 1) < 0.01


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("ncontrasts", [1, 2])
def test_cartesian_known_answers(ndim, ncontrasts):
    """Restates deepmr's tests/fft/test_nufft.py: integer coordinates take the
    width-1 path; a centred delta maps to all ones and back."""
    npix, ncoils = 4, 2
    print(f"Cartesian known answer ndim={ndim} ncontrasts={ncontrasts}", flush=True)
    nodes = np.arange(npix) - npix // 2
    mesh = np.meshgrid(*([nodes] * ndim), indexing="ij")
    coord = np.stack([m.ravel() for m in mesh], -1).astype(np.float32)[None]
    if ncontrasts > 1:
        coord = np.repeat(coord[None], ncontrasts, 0)
    lead = (ncoils,) if ncontrasts == 1 else (ncoils, ncontrasts)
    img = torch.zeros((*lead, *(npix,) * ndim), dtype=torch.complex64)
    img[(Ellipsis,) + (npix // 2,) * ndim] = 1.0
    k = tilenufft.fft.nufft(img, coord=torch.from_numpy(coord), device=DEV, width=12)
    assert k.device.type == "cpu"
    np.testing.assert_allclose(k.numpy(), np.ones((*lead, 1, npix**ndim)), rtol=1e-4, atol=1e-4)
    back = tilenufft.fft.nufft_adj(torch.ones((*lead, 1, npix**ndim), dtype=torch.complex64),
                                   coord=torch.from_numpy(coord), shape=[npix] * ndim, device=DEV,
                                   width=12)
    np.testing.assert_allclose(back.numpy(), img.numpy(), rtol=1e-4, atol=1e-4)


def test_plain_interpolator_matches_oracle():
    n = 40
    print("plain interpolator/gridding 2D", flush=True)
    rng = np.random.default_rng(11)
    coord = rng.uniform(0, n, (2, 5, 30, 2)).astype(np.float32)  # 2 frames, grid units
    interp = tilenufft._interp.plan_interpolator(coord, n, width=3, beta=2.5, device=DEV)
    oi = orc.plan_interpolator(coord, [n, n], [3, 3], [np.float32(2.5)] * 2)
    g = cnormal((3, 2, n, n), 12)
    s = tilenufft._interp.apply_interpolation(torch.from_numpy(g).to(DEV), interp).cpu().numpy()
    assert rel(s, orc.apply_interpolation(g, oi)) < RTOL
    ks = cnormal((3, 2, 5, 30), 13)
    gg = tilenufft._interp.apply_gridding(torch.from_numpy(ks).to(DEV), interp).cpu().numpy()
    assert rel(gg, orc.apply_gridding(ks, oi)) < RTOL


@pytest.mark.parametrize("ndim", [2, 3])
def test_interpolate_and_gridding_wrappers_match_oracle(ndim):
    """deepmr's one-call wrappers (plan + apply, _interp/__init__.py:15-156)
    with their default width 2 and beta 1 and with width 3: two frames, a
    coil batch, numpy in and a tensor on the input's device out, as deepmr's,
    against the oracle's plan + apply."""
    n = 24 if ndim == 2 else 12
    print(f"interpolate / gridding wrappers {ndim}D", flush=True)
    rng = np.random.default_rng(14 + ndim)
    coord = rng.uniform(0, n, (2, 4, 25, ndim)).astype(np.float32)
    g = cnormal((3, 2) + (n,) * ndim, 15)
    ks = cnormal((3, 2, 4, 25), 16)
    for kw in ({}, {"width": 3, "beta": 2.5}):
        w, b = kw.get("width", 2), kw.get("beta", 1.0)
        oi = orc.plan_interpolator(coord, [n] * ndim, [w] * ndim, [np.float32(b)] * ndim)
        s = tilenufft._interp.interpolate(g, torch.from_numpy(coord), device=DEV, **kw)
        assert s.device.type == "cpu" and tuple(s.shape) == (3, 2, 4, 25)
        assert rel(s.numpy(), orc.apply_interpolation(g, oi)) < RTOL
        gg = tilenufft._interp.gridding(ks, torch.from_numpy(coord), [n] * ndim, device=DEV, **kw)
        assert gg.device.type == "cpu" and tuple(gg.shape) == (3, 2) + (n,) * ndim
        assert rel(gg.numpy(), orc.apply_gridding(ks, oi)) < RTOL


def test_numpy_in_numpy_out_and_real_input():
    n = 16
    coord = testing.golden_angle_radial(8, 32, n)
    x = np.random.default_rng(0).standard_normal((n, n)).astype(np.float32)
    y = tilenufft.fft.nufft(x, coord, device=DEV, oversamp=2.0)
    assert isinstance(y, np.ndarray) and y.dtype == np.complex64
    assert rel(y, orc.nufft(x, coord, n, 4, 2.0)) < RTOL


@pytest.mark.parametrize("ndim, mode, itile", [(2, 1, None), (2, 2, None), (2, 2, (16, 16)),
                                               (2, 2, (7, 13)), (3, 2, None), (3, 2, (4, 8, 8)),
                                               (3, 1, None), (3, 2, (5, 3, 7))])
def test_interp_variants_match_oracle(ndim, mode, itile):
    # gathering and LDS-staged interpolation, including tiles that do not
    # divide the grid (partial edge tiles, regions wrapping around the grid)
    n = 36 if ndim == 2 else 14
    print(f"interp variant ndim={ndim} mode={mode} itile={itile}", flush=True)
    coord = testing.golden_angle_radial(24, 72, n) if ndim == 2 else testing.kooshball(50, 28, n)
    x = cnormal((3,) + (n,) * ndim, 31)
    plan = tilenufft.fft.plan_nufft(coord, n, 4, 1.5, device=DEV, interp_mode=mode, itile=itile)
    y = tilenufft.fft.apply_nufft(torch.from_numpy(x).to(DEV), plan).cpu().numpy()
    assert rel(y, orc.apply_nufft(x, orc.plan_nufft(coord, n, 4, 1.5))) < RTOL


def test_staged_interp_3d_same_bits_for_every_tiling():
    """3D staged interpolation takes each tile's samples in footprint-origin
    order, a segment of them per workgroup: whichever tile, segment, lane or
    coil chunk forms a sample's value, its taps are summed in the same order,
    so the k-space is the same bits for every tiling and chunk, and matches
    the oracle. A dense kooshball centre gives origin cells shared by many
    samples, segments splitting a tile and wrapped regions at the faces."""
    n = 16
    coord = np.concatenate([testing.kooshball(120, 32, n), np.random.default_rng(5).uniform(
        -n / 2, n / 2, (20, 32, 3)).astype(np.float32)])
    x = cnormal((5, n, n, n), 33)
    xt = torch.from_numpy(x).to(DEV)
    outs = []
    for itile, chunk in [((8, 8, 16), 2), ((8, 8, 8), 4), ((4, 8, 8), 1), ((4, 4, 4), 3)]:
        print(f"staged 3D interp itile={itile} chunk={chunk}", flush=True)
        plan = tilenufft.fft.plan_nufft(coord, n, 4, 2.0, device=DEV, interp_mode=2, itile=itile,
                                        interp_chunk=chunk)
        outs.append(tilenufft.fft.apply_nufft(xt, plan))
    torch.cuda.synchronize()
    for o in outs[1:]:
        assert torch.equal(o.view(torch.float32), outs[0].view(torch.float32))
    assert rel(outs[0].cpu().numpy(), orc.apply_nufft(x, orc.plan_nufft(coord, n, 4, 2.0))) < RTOL


def test_staged_interp_rows_longer_than_workgroup():
    """Interpolation tiles whose region rows (tile + W - 1 cells) are longer
    than the 256 lanes of a workgroup take the cell-by-cell region loads:
    the k-space matches the oracle and is the same bits as with the default
    16 x 16 tiles (a sample's sum does not depend on the tiling)."""
    n = 192
    print("staged interp rows longer than the workgroup", flush=True)
    coord = testing.golden_angle_radial(16, 128, n)
    x = cnormal((2, n, n), 34)
    xt = torch.from_numpy(x).to(DEV)
    wide = tilenufft.fft.plan_nufft(coord, n, 4, 1.5, device=DEV, interp_mode=2, itile=(4, 288))
    assert wide.os_shape == (288, 288)
    base = tilenufft.fft.plan_nufft(coord, n, 4, 1.5, device=DEV, interp_mode=2, itile=(16, 16))
    a = tilenufft.fft.apply_nufft(xt, wide)
    b = tilenufft.fft.apply_nufft(xt, base)
    torch.cuda.synchronize()
    assert torch.equal(a.view(torch.float32), b.view(torch.float32))
    assert rel(a.cpu().numpy(), orc.apply_nufft(x, orc.plan_nufft(coord, n, 4, 1.5))) < RTOL


@pytest.mark.parametrize("split", [-1, 3, 64, 500])
def test_grid_split_matches_oracle_and_is_deterministic(split):
    # hot tiles split across workgroups, partials folded in a fixed order
    n = 40
    print(f"gridding split={split}", flush=True)
    coord = testing.vd_spiral(12, 300, n)  # dense centre -> hot tiles
    k = cnormal((3, 12, 300), 41)
    plan = tilenufft.fft.plan_nufft(coord, n, 4, 2.0, device=DEV, grid_split=split)
    kt = torch.from_numpy(k).to(DEV)
    a = tilenufft.fft.apply_nufft_adj(kt, plan)
    b = tilenufft.fft.apply_nufft_adj(kt, plan)
    torch.cuda.synchronize()
    assert torch.equal(a.view(torch.float32), b.view(torch.float32))
    assert rel(a.cpu().numpy(), orc.apply_nufft_adj(k, orc.plan_nufft(coord, n, 4, 2.0))) < RTOL


def test_conjugate_view_inputs():
    """Lazy conjugate views (x.conj()) are materialised before the kernels."""
    n = 32
    print("conj-view inputs", flush=True)
    coord = testing.golden_angle_radial(16, 64, n)
    plan = tilenufft.fft.plan_nufft(coord, n, 4, 2.0, device=DEV)
    x = torch.from_numpy(cnormal((2, n, n), 21)).to(DEV)
    y1 = tilenufft.fft.apply_nufft(x.conj(), plan)
    y2 = tilenufft.fft.apply_nufft(x.conj().resolve_conj(), plan)
    assert torch.equal(y1, y2)
    k = torch.from_numpy(cnormal(tuple(y1.shape), 22)).to(DEV)
    a1 = tilenufft.fft.apply_nufft_adj(k.conj(), plan)
    a2 = tilenufft.fft.apply_nufft_adj(k.conj().resolve_conj(), plan)
    assert torch.equal(a1, a2)


GRID_CASES = [
    # (ndim, N, width, osf, coils, split)
    (2, 40, 4, 2.0, 1, 0),    # one coil: waves own row bands
    (2, 40, 4, 2.0, 3, 3),    # tiny split: every tile in parts
    (2, 40, 4, 2.0, 8, 64),
    (2, 24, 4, 2.0, 20, 0),   # 20 coils: 8-lane coil groups, idle lanes
    (2, 36, 6, 1.5, 2, 0),    # 36 taps
    (2, 32, 8, 2.0, 5, 128),  # 64 taps per plane (per-sample atomics only)
    (2, 30, 5, 2.0, 6, 0),    # 25 taps
    (3, 14, 4, 1.5, 4, 0),    # 4 z-plane passes of 16 taps
    (3, 16, 6, 2.0, 3, 64),   # 6 passes of 36 taps
    (3, 14, 3, 2.0, 2, 16),
    (3, 12, 2, 2.0, 9, 0),
]


@pytest.mark.parametrize("mode", [1, 2, 3])
@pytest.mark.parametrize("ndim, n, width, osf, ncoil, split", GRID_CASES)
def test_grid_modes_match_oracle_and_are_deterministic(mode, ndim, n, width, osf, ncoil, split):
    """The gridding variants (per-sample LDS atomics over origin-spread
    entries; register runs over origin-sorted entries; per-cell sums over the
    sparse gridding matrix) on dense-centre trajectories, split tiles, coil
    counts below and above the wave count and the coils per lane."""
    if mode == 2 and width * width > 36:
        pytest.skip("register-run gridding holds at most 36 taps per z plane")
    print(f"grid mode={mode} ndim={ndim} N={n} W={width} osf={osf} coils={ncoil} split={split}", flush=True)
    if ndim == 2:
        coord = testing.vd_spiral(10, 256, n)  # many samples share the centre origins
    else:
        coord = testing.kooshball(60, 32, n)
    k = cnormal((ncoil,) + coord.shape[:-1], 50 + width)
    w = np.random.default_rng(51).uniform(0.5, 1.5, coord.shape[:-1]).astype(np.float32)
    plan = tilenufft.fft.plan_nufft(coord, n, width, osf, device=DEV, grid_mode=mode, grid_split=split)
    kt = torch.from_numpy(k).to(DEV)
    wt = torch.from_numpy(w)
    a = tilenufft.fft.apply_nufft_adj(kt, plan, weight=wt)
    b = tilenufft.fft.apply_nufft_adj(kt, plan, weight=wt)
    torch.cuda.synchronize()
    assert torch.equal(a.view(torch.float32), b.view(torch.float32))
    ref = orc.apply_nufft_adj(k, orc.plan_nufft(coord, n, width, osf), weight=w)
    assert rel(a.cpu().numpy(), ref) < RTOL


@pytest.mark.parametrize("chunk", [1, 3, 8, 16])
def test_grid_matrix_coil_chunks(chunk):
    """Matrix gridding with 1..16 coils per item over 21 coils (records of 4,
    8 and 16 coils summed two coils per lane in 2D; partial last items hold
    zero coils that are not stored) and a density weight, against the
    oracle."""
    n = 32
    print(f"grid matrix chunk={chunk}", flush=True)
    coord = testing.vd_spiral(8, 200, n)
    k = cnormal((21,) + coord.shape[:-1], 53)
    w = np.random.default_rng(54).uniform(0.5, 1.5, coord.shape[:-1]).astype(np.float32)
    plan = tilenufft.fft.plan_nufft(coord, n, 4, 2.0, device=DEV, grid_mode=3, grid_chunk=chunk)
    a = tilenufft.fft.apply_nufft_adj(torch.from_numpy(k).to(DEV), plan, weight=torch.from_numpy(w))
    ref = orc.apply_nufft_adj(k, orc.plan_nufft(coord, n, 4, 2.0), weight=w)
    assert rel(a.cpu().numpy(), ref) < RTOL


def test_grid_matrix_multiframe_plan_and_empty_frame_rows():
    """One plan of several frames (per-frame row offsets into one matrix),
    with repeated samples on the grid edge (taps that wrap), against the
    oracle."""
    n = 24
    print("grid matrix multi-frame", flush=True)
    rng = np.random.default_rng(55)
    coord = rng.uniform(-n / 2, n / 2, (3, 5, 40, 2)).astype(np.float32)
    coord[1, 0, :5] = [[-n / 2, -n / 2]] * 5  # a duplicated edge sample
    k = cnormal((2, 3, 5, 40), 56)
    plan = tilenufft.fft.plan_nufft(coord, n, 4, 1.5, device=DEV, grid_mode=3)
    a = tilenufft.fft.apply_nufft_adj(torch.from_numpy(k).to(DEV), plan)
    ref = orc.apply_nufft_adj(k, orc.plan_nufft(coord, n, 4, 1.5))
    assert rel(a.cpu().numpy(), ref) < RTOL


@pytest.mark.parametrize("ncoil, chunk", [(16, 0), (8, 0), (3, 0), (2, 1)])
def test_grid_matrix_long_rows_in_parts(ncoil, chunk):
    """Cells with more entries than one workgroup part (3000 samples on one
    spot next to a uniform background): the rows are summed in parts whose
    partials are folded in order; coil-minor (chunks of 16, 8 and 3 coils)
    and lane-per-cell (1 coil per item) kernels, against the oracle, twice."""
    n = 16
    print(f"grid matrix long rows coils={ncoil} chunk={chunk}", flush=True)
    rng = np.random.default_rng(57)
    bg = rng.uniform(-n / 2, n / 2, (1, 400, 2))
    hot = np.tile(np.array([[[0.3, -0.2]]]), (1, 3000, 1)) + rng.normal(0, 0.05, (1, 3000, 2))
    coord = np.concatenate([bg, hot], 1).astype(np.float32)
    k = cnormal((ncoil,) + coord.shape[:-1], 58)
    plan = tilenufft.fft.plan_nufft(coord, n, 4, 2.0, device=DEV, grid_mode=3, grid_chunk=chunk)
    kt = torch.from_numpy(k).to(DEV)
    a = tilenufft.fft.apply_nufft_adj(kt, plan)
    b = tilenufft.fft.apply_nufft_adj(kt, plan)
    torch.cuda.synchronize()
    assert torch.equal(a.view(torch.float32), b.view(torch.float32))
    ref = orc.apply_nufft_adj(k, orc.plan_nufft(coord, n, 4, 2.0))
    assert rel(a.cpu().numpy(), ref) < RTOL


@pytest.mark.parametrize("split", [0, 64])
def test_grid_matrix_sums_do_not_depend_on_record_width(split):
    """A coil's gridded values are bitwise the same whether it is gridded in
    a 16-coil record, a 4-coil record (3 coils) or in place (one coil per
    item): short rows are summed in entry order and long rows (a 3000-sample
    hot spot, rows above 64 entries with split 64) by a fixed partition
    whose group sums are added in order, for every record width."""
    n = 16
    print(f"grid matrix record-width independence split={split}", flush=True)
    rng = np.random.default_rng(59)
    bg = rng.uniform(-n / 2, n / 2, (1, 400, 2))
    hot = np.tile(np.array([[[0.3, -0.2]]]), (1, 3000, 1)) + rng.normal(0, 0.05, (1, 3000, 2))
    coord = np.concatenate([bg, hot], 1).astype(np.float32)
    kt = torch.from_numpy(cnormal((16,) + coord.shape[:-1], 60)).to(DEV)
    plan = tilenufft.fft.plan_nufft(coord, n, 4, 2.0, device=DEV, grid_mode=3, grid_split=split)
    plan1 = tilenufft.fft.plan_nufft(coord, n, 4, 2.0, device=DEV, grid_mode=3, grid_split=split, grid_chunk=1)
    a16 = tilenufft.fft.apply_nufft_adj(kt, plan)[:3]
    a3 = tilenufft.fft.apply_nufft_adj(kt[:3].contiguous(), plan)
    a1 = tilenufft.fft.apply_nufft_adj(kt[:3].contiguous(), plan1)
    torch.cuda.synchronize()
    assert torch.equal(a16.view(torch.float32), a3.view(torch.float32))
    assert torch.equal(a16.view(torch.float32), a1.view(torch.float32))


@pytest.mark.parametrize("chunk, split", [(6, 32), (16, 32), (16, 0)])
def test_grid_matrix_3d_records_and_parts(chunk, split):
    """3D matrix gridding over 11 coils: 8-coil records (one coil per lane in
    3D) and 16-coil records (two coils per lane), with the dense kooshball
    centre summed in parts when rows exceed 32 entries, against the oracle,
    twice (bitwise)."""
    n = 16
    print(f"grid matrix 3D chunk={chunk} split={split}", flush=True)
    coord = testing.kooshball(300, 32, n)
    k = cnormal((11,) + coord.shape[:-1], 59)
    plan = tilenufft.fft.plan_nufft(coord, n, 4, 1.5, device=DEV, grid_mode=3, grid_chunk=chunk,
                                    grid_split=split)
    kt = torch.from_numpy(k).to(DEV)
    a = tilenufft.fft.apply_nufft_adj(kt, plan)
    b = tilenufft.fft.apply_nufft_adj(kt, plan)
    torch.cuda.synchronize()
    assert torch.equal(a.view(torch.float32), b.view(torch.float32))
    ref = orc.apply_nufft_adj(k, orc.plan_nufft(coord, n, 4, 1.5))
    assert rel(a.cpu().numpy(), ref) < RTOL


def test_grid_runs_refuses_wide_kernels():
    coord = testing.golden_angle_radial(12, 64, 24)
    with pytest.raises(RuntimeError, match="taps"):
        tilenufft.fft.plan_nufft(coord, 24, 7, 2.0, device=DEV, grid_mode=2)
    plan = tilenufft.fft.plan_nufft(coord, 24, 7, 2.0, device=DEV)  # automatic: the gridding matrix (49 taps)
    assert plan.interpolator.grid_mode_effective() == 3
    k = torch.from_numpy(cnormal((1,) + coord.shape[:-1], 52)).to(DEV)
    ref = orc.apply_nufft_adj(k.cpu().numpy(), orc.plan_nufft(coord, 24, 7, 2.0))
    assert rel(tilenufft.fft.apply_nufft_adj(k, plan).cpu().numpy(), ref) < RTOL


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_empty_trajectory_plan(mode):
    """A plan with no samples (every gridding variant): the adjoint writes an
    all-zero grid and image, the forward an empty k-space, through the plan
    and through a batched launch."""
    print(f"empty trajectory grid_mode={mode}", flush=True)
    coord = np.zeros((0, 8, 2), np.float32)  # (views, samples, 2) with no views
    plan = tilenufft.fft.plan_nufft(coord, 24, 4, 2.0, device=DEV, grid_mode=mode)
    k = torch.zeros((3, 0, 8), dtype=torch.complex64, device=DEV)
    img = tilenufft.fft.apply_nufft_adj(k, plan)
    assert img.shape == (3, 24, 24) and not torch.isnan(img.real).any() and float(img.abs().max()) == 0.0
    y = tilenufft.fft.apply_nufft(torch.ones((3, 24, 24), dtype=torch.complex64, device=DEV), plan)
    assert y.shape == (3, 0, 8)
    from tilenufft.batch import BatchedNUFFT

    b = BatchedNUFFT([plan, plan], 2)
    b.grid.fill_(float("nan"))
    b.gridding(b.grid)
    torch.cuda.synchronize()
    assert float(b.grid.abs().max()) == 0.0


@pytest.mark.parametrize("grid_mode", [0, 1, 2, 3])
@pytest.mark.parametrize("osf", [2.0, 1.25])
def test_nufft1d_matches_oracle(grid_mode, osf):
    """1D NUFFT (deepmr's _interpolate1 / _gridding1 path, interp.py:125,
    grid.py:182): two frames of 60 random samples, three coils, every
    gridding variant, forward and adjoint against the oracle."""
    n = 40
    print(f"1D NUFFT grid_mode={grid_mode} osf={osf}", flush=True)
    rng = np.random.default_rng(21)
    coord = rng.uniform(-n / 2, n / 2, (2, 1, 60, 1)).astype(np.float32)
    plan = tilenufft.fft.plan_nufft(coord, n, 4, osf, device=DEV, grid_mode=grid_mode)
    assert plan.ndim == 1 and plan.shape == (n,) and plan.os_shape == (int(np.ceil(osf * n)),)
    oplan = orc.plan_nufft(coord, n, 4, osf)
    x = cnormal((3, 2, n), 22)
    y = tilenufft.fft.apply_nufft(torch.from_numpy(x).to(DEV), plan).cpu().numpy()
    assert y.shape == (3, 2, 1, 60)
    assert rel(y, orc.apply_nufft(x, oplan)) < RTOL
    k = cnormal(y.shape, 23)
    xa = tilenufft.fft.apply_nufft_adj(torch.from_numpy(k).to(DEV), plan).cpu().numpy()
    assert xa.shape == (3, 2, n)
    assert rel(xa, orc.apply_nufft_adj(k, oplan)) < RTOL


@pytest.mark.parametrize("ncontrasts, ncoils, nslices", [(1, 1, 1), (1, 2, 2), (2, 1, 2), (2, 2, 1)])
def test_known_answers_1d(ncontrasts, ncoils, nslices):
    """Restates deepmr's tests/fft/test_nufft.py test_nufft1 (:28-62): a
    centred delta maps to all ones and back on integer coordinates."""
    npix = 4
    print(f"1D known answer ncontrasts={ncontrasts}", flush=True)
    coord = (np.arange(npix) - npix // 2).astype(np.float32)[None, :, None]
    if ncontrasts > 1:
        coord = np.repeat(coord[None], ncontrasts, 0)
    lead = (nslices, ncoils) if ncontrasts == 1 else (nslices, ncoils, ncontrasts)
    img = torch.zeros((*lead, npix), dtype=torch.complex64)
    img[..., npix // 2] = 1.0
    k = tilenufft.fft.nufft(img, coord=torch.from_numpy(coord), device=DEV, width=12)
    np.testing.assert_allclose(k.numpy(), np.ones((*lead, 1, npix)), rtol=1e-4, atol=1e-4)
    back = tilenufft.fft.nufft_adj(torch.ones((*lead, 1, npix), dtype=torch.complex64),
                                   coord=torch.from_numpy(coord), shape=[npix], device=DEV, width=12)
    np.testing.assert_allclose(back.numpy(), img.numpy(), rtol=1e-4, atol=1e-4)


def test_lowrank_1d_matches_oracle():
    """1D subspace NUFFT (deepmr's _interpolate_lowrank1 / _gridding_lowrank1,
    interp.py:231, grid.py:288) against the oracle, and the reference's
    identity-basis known answer (tests/fft/test_nufft.py:65-97)."""
    nf, nc, n = 4, 2, 32
    print("1D low rank", flush=True)
    rng = np.random.default_rng(24)
    coord = rng.uniform(-n / 2, n / 2, (nf, 1, 30, 1)).astype(np.float32)
    plan = tilenufft.fft.plan_nufft(coord, n, 4, 2.0, device=DEV)
    oplan = orc.plan_nufft(coord, n, 4, 2.0)
    basis = cnormal((nc, nf), 25)
    ba = basis.conj().T.copy()
    coef = cnormal((3, nc, n), 26)
    y = tilenufft.fft.apply_nufft(torch.from_numpy(coef).to(DEV), plan, basis_adjoint=torch.from_numpy(ba))
    assert rel(y.cpu().numpy(), orc.apply_nufft(coef, oplan, basis_adjoint=ba)) < RTOL
    k = cnormal((3, nf, 1, 30), 27)
    g = tilenufft.fft.apply_nufft_adj(torch.from_numpy(k).to(DEV), plan, basis=torch.from_numpy(basis))
    assert rel(g.cpu().numpy(), orc.apply_nufft_adj(k, oplan, basis=basis)) < RTOL
    npix = 4
    cc = np.repeat((np.arange(npix) - npix // 2).astype(np.float32)[None, None, :, None], 3, 0)
    img = torch.zeros((2, 2, 3, npix), dtype=torch.complex64)
    img[..., npix // 2] = 1.0
    kk = tilenufft.fft.nufft(img, coord=torch.from_numpy(cc), device=DEV, width=12,
                             basis_adjoint=torch.eye(3, dtype=torch.complex64))
    np.testing.assert_allclose(kk.numpy(), np.ones((2, 2, 3, 1, npix)), rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("ndim, n, width", [(2, 24, 12), (2, 20, 16), (3, 12, 10)])
def test_wide_kernels_match_oracle(ndim, n, width):
    """Kernel widths above 8 (deepmr accepts any width): the gathering
    interpolation and the matrix gridding, forward and adjoint, plus the
    low-rank pair in 2D, against the oracle; the tile gridding variants
    refuse them with an error."""
    print(f"wide kernel ndim={ndim} N={n} W={width}", flush=True)
    rng = np.random.default_rng(31 + width)
    coord = rng.uniform(-n / 2, n / 2, (2, 1, 50, ndim)).astype(np.float32)
    plan = tilenufft.fft.plan_nufft(coord, n, width, 2.0, device=DEV)
    oplan = orc.plan_nufft(coord, n, width, 2.0)
    x = cnormal((2, 2) + (n,) * ndim, 32)
    y = tilenufft.fft.apply_nufft(torch.from_numpy(x).to(DEV), plan).cpu().numpy()
    assert rel(y, orc.apply_nufft(x, oplan)) < RTOL
    k = cnormal(y.shape, 33)
    xa = tilenufft.fft.apply_nufft_adj(torch.from_numpy(k).to(DEV), plan).cpu().numpy()
    assert rel(xa, orc.apply_nufft_adj(k, oplan)) < RTOL
    if ndim == 2:
        basis = cnormal((1, 2), 34)  # (ncoeff, nframes)
        ba = basis.conj().T.copy()
        coef = cnormal((2, 1, n, n), 35)
        yl = tilenufft.fft.apply_nufft(torch.from_numpy(coef).to(DEV), plan, basis_adjoint=torch.from_numpy(ba))
        assert rel(yl.cpu().numpy(), orc.apply_nufft(coef, oplan, basis_adjoint=ba)) < RTOL
        gl = tilenufft.fft.apply_nufft_adj(torch.from_numpy(k).to(DEV), plan, basis=torch.from_numpy(basis))
        assert rel(gl.cpu().numpy(), orc.apply_nufft_adj(k, oplan, basis=basis)) < RTOL
    for mode in (1, 2):
        with pytest.raises(RuntimeError):
            tilenufft.fft.plan_nufft(coord, n, width, 2.0, device=DEV, grid_mode=mode)
