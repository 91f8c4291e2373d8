"""Pins for the oracle's plan arithmetic (CPU only).

Each check is chosen so that a plausible mistake in the restatement (a wrong
polynomial coefficient, an off-by-one tap origin, a dropped wrap, a wrong
scale) fails it, using a source independent of the restated formula:
scipy's I0, brute-force tap windows, and the on-grid interpolation identity.
"""

import math

import numpy as np
import pytest
from scipy.special import i0

from oracle import ndft
from oracle import nufft_oracle as orc


@pytest.mark.parametrize("beta", [2.0, 6.99666, 7.89229, 8.99615, 12.0, 16.0])
def test_kb_polynomial_matches_scipy_i0(beta):
    # A&S 9.8.1 / 9.8.2 are accurate to ~1.6e-7 / ~1.9e-7 (relative <= 5e-7)
    u = np.linspace(-1, 1, 8001)
    ref = i0(beta * np.sqrt(1 - u**2))
    got = orc.kaiser_bessel(u, beta)
    assert np.max(np.abs(got / ref - 1)) < 5e-7
    lib = orc.load_loops()
    got_c = np.array([lib.orc_kaiser_bessel(float(v), beta) for v in u[::40]])
    assert np.max(np.abs(got_c / ref[::40] - 1)) < 5e-7


def test_kb_support_and_edges():
    assert orc.kaiser_bessel(np.array([1.0000001, -1.5, 2.0]), 8.0).tolist() == [0.0, 0.0, 0.0]
    # KB(+-1) = I0(0) = 1 (the reference does not zero the edge taps)
    assert np.allclose(orc.kaiser_bessel(np.array([-1.0, 1.0]), 8.0), 1.0)


def test_beta_values_and_optimality():
    # Beatty et al. 2005: beta = pi*sqrt((W/a*(a-1/2))^2 - 0.8)
    np.testing.assert_allclose(orc.kb_beta([4], [2.0]), [8.99615], atol=5e-5)
    np.testing.assert_allclose(orc.kb_beta([4], [1.5]), [7.89229], atol=5e-5)
    np.testing.assert_allclose(orc.kb_beta([4], [1.25]), [6.99666], atol=5e-5)
    # the formula's beta must beat detuned values on the NDFT error
    n = 24
    rng = np.random.default_rng(0)
    coord = rng.uniform(-n / 2, n / 2, (400, 2)).astype(np.float32)
    x = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))).astype(np.complex64)
    exact = ndft.ndft(x, coord, (n, n))

    def err(beta_scale):
        plan = orc.plan_nufft(coord, n, 4, 2.0)
        beta = tuple(np.float32(b * beta_scale) for b in plan["beta"])
        scaled = plan["scaled_coord"]
        plan["interpolator"] = orc.plan_interpolator(scaled, plan["os_shape"], plan["width"], beta)
        plan["beta"] = beta
        y = orc.apply_nufft(x, plan)
        return np.linalg.norm(y - exact) / np.linalg.norm(exact)

    e0 = err(1.0)
    assert e0 < 2e-3
    assert e0 < err(0.85) and e0 < err(1.15)


def _brute_taps(g, width, grid, beta):
    """Integers k with (k - g)/(W/2) >= -1, first W of them, from a wide scan."""
    ks = np.arange(math.floor(g) - 2 * width, math.floor(g) + 2 * width + 1)
    ks = ks[(ks - g) / (width / 2) >= -1][:width]
    u = (ks - g) / (width / 2)
    vals = i0(beta * np.sqrt(np.clip(1 - u**2, 0, None)))
    return np.mod(ks, grid), vals


@pytest.mark.parametrize("width", [2, 3, 4, 5, 6])
def test_tap_tables_brute_force(width):
    grid, beta = 17, np.float32(7.5)
    rng = np.random.default_rng(width)
    g = np.concatenate([rng.uniform(-3, grid + 3, 200), np.arange(-2, grid + 2, 0.5),
                        [0.0, grid - 1e-3, grid - 0.5, 1e-6]]).astype(np.float32)
    coord = g[:, None]  # 1 axis, x
    value = np.zeros((g.size, width), np.float32)
    index = np.zeros((g.size, width), np.int32)
    orc.load_loops().orc_prepare_interpolator(orc._ptr(value), orc._ptr(index),
                                              orc._ptr(np.ascontiguousarray(coord[:, 0])),
                                              g.size, width, beta, grid)
    for i, gi in enumerate(g.astype(np.float64)):
        idx, val = _brute_taps(gi, width, grid, float(beta))
        assert index[i].tolist() == idx.tolist(), (gi, index[i], idx)
        np.testing.assert_allclose(value[i], val, rtol=5e-7)


@pytest.mark.parametrize("width, osf", [(4, 2.0), (4, 1.25), (3, 1.5), (6, 2.0)])
def test_scale_on_grid_identity(width, osf):
    # a sample on a grid node interpolates an all-ones grid to exactly
    # sum(taps)/scale = 1; a wrong scale or tap origin breaks this
    n = 20
    grid = int(np.ceil(np.float32(osf) * n))
    g = np.arange(grid // 4, 3 * grid // 4, dtype=np.float32)
    c = ((g - grid // 2) * np.float32(n / grid)).astype(np.float32)
    coord = np.stack(np.meshgrid(c, c, indexing="ij"), -1).reshape(-1, 2)
    # one irrational sample keeps the plan off the Cartesian width-1 path
    coord = np.concatenate([coord, np.array([[0.1234567, -0.7654321]], np.float32)])
    plan = orc.plan_nufft(coord, n, width, osf)
    assert plan["width"] == (width, width)
    interp = plan["interpolator"]
    g_ones = np.ones((1, 1, *plan["os_shape"]), np.complex64)
    y = orc.interpolate_raw(g_ones, interp)[0, 0] / np.float32(interp["scale"])
    on_node = np.all(np.mod(plan["scaled_coord"].reshape(-1, 2), 1.0) == 0, axis=-1)
    assert on_node.sum() > 10
    np.testing.assert_allclose(y[on_node], 1.0, rtol=2e-6)
    # closed form with an independent I0
    start = math.ceil(-width / 2)
    u = (start + np.arange(width)) / (width / 2)
    s1 = i0(float(plan["beta"][0]) * np.sqrt(1 - u**2)).sum()
    np.testing.assert_allclose(interp["scale"], s1**2, rtol=1e-6)


def test_cartesian_axis_quirk():
    # shape indexed (z, y, x) against coord[..., ax] in (x, y, z) order
    coord = np.zeros((5, 2), np.float32)
    coord[:, 0] = np.arange(5) * 0.5  # x: half-integers
    coord[:, 1] = np.arange(5)  # y: integers
    # shape (ny=3, nx=2). A natural pairing would test x with nx=2 (2*0.5k
    # integer) and y with ny=3 (integer): both Cartesian. deepmr tests
    # coord[..., 0] (x) with shape[0] = ny = 3 -> 1.5k, not integer, and
    # coord[..., 1] (y) with shape[1] = nx = 2 -> integer; reversed to (y, x).
    flags = orc.cartesian_axes(coord, np.array([3, 2], np.int16))
    assert flags.tolist() == [True, False]
