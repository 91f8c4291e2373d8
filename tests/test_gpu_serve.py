"""Continuous-batching server (GPU).

- a request's result is the same alone and in a mix of other sizes / coil
  counts / trajectories, and equal to the plain one-request-at-a-time NUFFT
  path: the SENSE combination and CG update are shared native kernels with
  a fixed reduction order, and the NUFFT stages of a request do not depend on
  the other requests of a launch;
- the served reconstruction agrees with CG-SENSE over the oracle NUFFT as
  closely as the oracle's own float32 result agrees with its float64 one;
- the same at full size (N 320, 32 coils, ~200k samples, 10 iterations);
- admission respects the HBM budget and the slot limit, with requests
  joining while others are mid-solve, and the budget figure of a request
  covers the device memory its plan takes;
- a malformed, an oversized and a cancelled request fail alone, and a plan
  with a Cartesian axis (another oversampled grid) is served in its own group.
Failure conditions are forced from the host; no kernel faults.
"""

import copy

import numpy as np
import pytest
import torch

from tilenufft import fft as tfft
from tilenufft import serve, testing

from oracle import autograd_ref, nufft64
from oracle import nufft_oracle as orc

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def rel(a, b):
    a = torch.as_tensor(a).cpu().to(torch.complex128)
    b = torch.as_tensor(b).cpu().to(torch.complex128)
    return float(torch.linalg.norm(a - b) / torch.linalg.norm(b))


def small_stream(n=6, seed=0, sizes=(64, 128, 192), niter=4):
    return serve.make_stream(n, rate_hz=0, seed=seed, sizes=sizes, coils=(2, 6),
                             samples=(3000, 12000), niter=niter)


@pytest.mark.parametrize("max_active, admission", [(32, "fifo"), (3, "sjf")])
def test_alone_equals_mix_equals_plain(max_active, admission):
    print(f"server: alone vs mix vs plain ({max_active} active, {admission})", flush=True)
    reqs = small_stream(8)
    alone = copy.deepcopy(reqs[2])
    srv = serve.ReconServer(DEV, hbm_budget=4 << 30)
    srv.run([alone], realtime=False)
    srv.close()
    srv = serve.ReconServer(DEV, hbm_budget=4 << 30, max_active=max_active, admission=admission)
    srv.run(reqs, realtime=False)
    srv.close()
    assert all(r.error is None for r in reqs), [r.error for r in reqs]
    assert rel(reqs[2].result, alone.result) < 1e-6
    plain = copy.deepcopy(reqs)
    serve.run_plain(plain, DEV, realtime=False)
    for a, b in zip(reqs, plain):
        assert b.error is None, b.error
        assert rel(a.result, b.result) < 1e-6, a.rid


def test_served_matches_oracle_cgsense():
    """Reduced size: N 64, 4 coils, 40 radial spokes x 128, 10 iterations."""
    print("server: served CG-SENSE vs the oracle's (float32 and float64)", flush=True)
    n, nc, niter = 64, 4, 10
    coord = testing.golden_angle_radial(40, 128, n)
    smaps = testing.gaussian_coil_maps(nc, (n, n), seed=3)
    y = testing.complex_normal((nc, 40, 128), seed=5)
    req = serve.ReconRequest(rid=0, coord=coord, kspace=y, smaps=torch.from_numpy(smaps), shape=n, niter=niter)
    other = small_stream(3, seed=4)
    srv = serve.ReconServer(DEV, hbm_budget=4 << 30)
    srv.run([req] + other, realtime=False)
    srv.close()
    assert req.error is None, req.error
    x = req.result.cpu().numpy()
    plan = orc.plan_nufft(coord, n, 4, 1.25)
    with torch.no_grad():
        x32 = autograd_ref.cgsense(y, torch.from_numpy(smaps), niter,
                                   lambda v: torch.from_numpy(orc.apply_nufft(v.numpy(), plan)),
                                   lambda k: torch.from_numpy(orc.apply_nufft_adj(k.numpy(), plan))).numpy()
    pl64 = nufft64.plan(coord, n, 4, 1.25)
    x64 = autograd_ref.cgsense64(y.numpy(), smaps, niter, pl64)
    e_srv, e_orc = rel(x, x64), rel(x32, x64)
    print(f"  served vs fp64 {e_srv:.2e}, oracle fp32 vs fp64 {e_orc:.2e}, served vs oracle fp32 {rel(x, x32):.2e}",
          flush=True)
    assert e_srv <= 2.0 * e_orc + 1e-6
    assert rel(x, x32) <= 2.0 * e_orc + 1e-6


def test_full_size_stream_alone_mix_plain():
    """N 320 with 32 coils and ~200k samples per coil among smaller requests."""
    print("server: full-size requests, alone vs mix vs plain", flush=True)
    reqs = serve.make_stream(5, rate_hz=0, seed=11, sizes=(128, 256, 320), coils=(4, 32),
                             samples=(10_000, 200_000), niter=10)
    big = serve.make_stream(1, rate_hz=0, seed=12, sizes=(320,), coils=(32, 32),
                            samples=(195_000, 200_000), niter=10)[0]
    big.rid = 99
    reqs.append(big)
    alone = copy.deepcopy(big)
    srv = serve.ReconServer(DEV, hbm_budget=32 << 30)
    srv.run([alone], realtime=False)
    srv.close()
    srv = serve.ReconServer(DEV, hbm_budget=32 << 30, max_active=32)  # all six in one mix
    srv.run(reqs, realtime=False)
    srv.close()
    assert all(r.error is None for r in reqs), [r.error for r in reqs]
    assert rel(big.result, alone.result) < 1e-6
    plain = copy.deepcopy(reqs)
    serve.run_plain(plain, DEV, realtime=False)
    worst = max(rel(a.result, b.result) for a, b in zip(reqs, plain))
    print(f"  max served vs plain {worst:.2e}", flush=True)
    assert worst < 1e-5


def test_budget_and_slots_limit_concurrency():
    print("server: admission by HBM budget / slots", flush=True)
    base = small_stream(1, seed=1)[0]
    reqs = []
    for i in range(6):  # equal sizes: exactly two fit the budget below
        r = copy.deepcopy(base)
        r.rid = i
        reqs.append(r)
    plan = tfft.plan_nufft(base.coord, base.shape, 4, 1.25, device=DEV)
    need = serve.request_bytes(base, plan)
    srv = serve.ReconServer(DEV, hbm_budget=int(2.5 * need), max_active=8)
    peak = []
    srv.run(reqs, realtime=False, on_step=lambda s: peak.append((len(s.active), s.reserved)))
    srv.close()
    assert all(r.error is None for r in reqs)
    assert max(a for a, _ in peak) == 2
    assert max(b for _, b in peak) <= srv.budget
    assert srv.stats["deferred_for_memory"] > 0
    # requests joined while others were active (continuous batching, not a queue)
    mixed = copy.deepcopy(small_stream(8, seed=1))
    for r in mixed:
        r.niter = 2 + r.rid % 5  # unequal lengths: slots free up at different steps
    srv2 = serve.ReconServer(DEV, hbm_budget=8 << 30, max_active=3)
    peak2 = []
    srv2.run(mixed, realtime=False, on_step=lambda s: peak2.append(len(s.active)))
    srv2.close()
    assert max(peak2) <= 3
    joined_mid = sum(1 for r in mixed if any(o.t_admit < r.t_admit < o.t_done for o in mixed if o is not r))
    assert joined_mid > 0


def test_request_bytes_cover_plan_memory():
    """The plan part of a request's budget is the library's count of its
    device memory, which must match what creating the plan takes from the
    device (free-memory delta, 2 MiB allocation granularity)."""
    print("server: budget vs device memory of a plan", flush=True)
    req = serve.make_stream(1, rate_hz=0, seed=21, sizes=(256,), coils=(16, 16),
                            samples=(150_000, 150_000))[0]
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    free0 = torch.cuda.mem_get_info()[0]
    plan = tfft.plan_nufft(req.coord, req.shape, 4, 1.25, device=DEV)
    torch.cuda.synchronize()
    torch.cuda.empty_cache()  # the coordinate upload's block goes back to the device
    free1 = torch.cuda.mem_get_info()[0]
    counted = plan.interpolator.device_bytes()
    taken = free0 - free1
    print(f"  plan bytes counted {counted}, device memory taken {taken}", flush=True)
    assert abs(counted - taken) <= 0.1 * taken + (16 << 20)
    assert serve.request_bytes(req, plan) > counted


def test_cartesian_axis_request_gets_its_own_group():
    """An on-grid x axis makes the plan width-1 with G = N on that axis; it
    is grouped apart from the oversampled plans and still equals plain."""
    print("server: Cartesian-axis request", flush=True)
    reqs = small_stream(3, seed=6, sizes=(64,))
    cart = copy.deepcopy(reqs[0])
    cart.rid = 50
    ny, nx = 48, 64
    ky, kx = np.meshgrid(np.linspace(-31.5, 31.5, ny, dtype=np.float32), np.arange(-32, 32, dtype=np.float32),
                         indexing="ij")
    cart.coord = np.stack([kx, ky + 0.25], -1).astype(np.float32)  # x on the grid, y off it
    cart.kspace = testing.complex_normal((cart.ncoils, ny, nx), seed=2)
    plan = tfft.plan_nufft(cart.coord, 64, 4, 1.25, device=DEV)
    assert plan.width[-1] == 1 and plan.os_shape[-1] == 64 and plan.os_shape[0] == 80
    reqs.append(cart)
    srv = serve.ReconServer(DEV, hbm_budget=4 << 30)
    srv.run(reqs, realtime=False)
    srv.close()
    assert all(r.error is None for r in reqs), [r.error for r in reqs]
    plain = copy.deepcopy(reqs)
    serve.run_plain(plain, DEV, realtime=False)
    for a, b in zip(reqs, plain):
        assert rel(a.result, b.result) < 1e-6, a.rid


def test_failures_stay_isolated():
    print("server: malformed / oversized / cancelled / empty requests", flush=True)
    reqs = small_stream(5, seed=2)
    ref = copy.deepcopy(reqs)
    srv = serve.ReconServer(DEV)
    srv.run(ref, realtime=False)
    srv.close()
    bad = copy.deepcopy(reqs[0])
    bad.rid = 100
    bad.coord = bad.coord.copy()
    bad.coord[0, 0, 0] = np.nan
    wrong = copy.deepcopy(reqs[1])
    wrong.rid = 101
    wrong.smaps = wrong.smaps[:, :-1]
    huge = copy.deepcopy(reqs[2])
    huge.rid = 102
    cancel = copy.deepcopy(reqs[3])
    cancel.rid = 103
    cancel.cancel()
    empty = copy.deepcopy(reqs[4])
    empty.rid = 104
    empty.coord = np.zeros((0, 2), np.float32)
    empty.kspace = torch.zeros((empty.ncoils, 0), dtype=torch.complex64)
    budget = int(1.5 * max(serve.estimate_request_bytes(r.shape, r.ncoils, r.npts) for r in reqs))
    huge.kspace = torch.zeros((96, huge.npts), dtype=torch.complex64)
    huge.smaps = torch.zeros((96, huge.shape, huge.shape), dtype=torch.complex64)
    assert serve.estimate_request_bytes(huge.shape, huge.ncoils, huge.npts) > budget
    srv = serve.ReconServer(DEV, hbm_budget=budget)
    allreq = reqs + [bad, wrong, huge, cancel, empty]
    srv.run(allreq, realtime=False)
    srv.close()
    assert "NaN" in bad.error
    assert "smaps" in wrong.error
    assert "budget" in huge.error
    assert cancel.error == "cancelled"
    assert "no samples" in empty.error
    for a, b in zip(reqs, ref):
        assert a.error is None
        assert rel(a.result, b.result) < 1e-6


def test_latency_stats_and_realtime_stream():
    reqs = serve.make_stream(6, rate_hz=50.0, seed=3, sizes=(64, 96), coils=(2, 4),
                             samples=(2000, 6000), niter=3)
    srv = serve.ReconServer(DEV)
    srv.run(reqs, realtime=True)
    srv.close()
    st = serve.latency_stats(reqs)
    assert st["completed"] == 6 and st["p50"] > 0 and st["p99"] >= st["p50"]
    assert all(r.t_done >= r.t_admit >= r.arrival - 1e-3 for r in reqs)
