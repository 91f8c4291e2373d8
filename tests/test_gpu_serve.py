"""Continuous-batching server (GPU).

- a request's result is the same alone and in a mix of other sizes / coil
  counts / trajectories (relative L2 <= 1e-6: interpolation and gridding are
  per-request bitwise, the batched FFT may change the last bits), and equals
  the plain one-request-at-a-time NUFFT path (<= 1e-5);
- admission respects the HBM budget and the slot limit, with requests
  joining while others are mid-solve;
- a malformed, an oversized and a cancelled request fail alone.
Failure conditions are forced from the host; no kernel faults.
"""

import copy

import numpy as np
import pytest
import torch

from tilenufft import serve

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def rel(a, b):
    return float(torch.linalg.norm(a.cpu() - b.cpu()) / torch.linalg.norm(b.cpu()))


def small_stream(n=6, seed=0):
    return serve.make_stream(n, rate_hz=0, seed=seed, sizes=(64, 96, 128), coils=(2, 6),
                             samples=(3000, 12000), niter=4)


def test_alone_equals_mix_equals_plain():
    print("server: alone vs mix vs plain", flush=True)
    reqs = small_stream()
    alone = copy.deepcopy(reqs[2])
    serve.ReconServer(DEV, hbm_budget=4 << 30).run([alone], realtime=False)
    srv = serve.ReconServer(DEV, hbm_budget=4 << 30)
    srv.run(reqs, realtime=False)
    assert all(r.error is None for r in reqs), [r.error for r in reqs]
    assert rel(reqs[2].result, alone.result) < 1e-6
    plain = copy.deepcopy(reqs)
    serve.run_plain(plain, DEV, realtime=False)
    for a, b in zip(reqs, plain):
        assert rel(a.result, b.result) < 1e-5, a.rid


def test_budget_and_slots_limit_concurrency():
    print("server: admission by HBM budget / slots", flush=True)
    base = small_stream(1, seed=1)[0]
    reqs = []
    for i in range(6):  # equal sizes: exactly two fit the budget below
        r = copy.deepcopy(base)
        r.rid = i
        reqs.append(r)
    need = serve.estimate_request_bytes(base.shape, base.ncoils, base.npts)
    srv = serve.ReconServer(DEV, hbm_budget=int(2.5 * need), max_active=8)
    peak = []
    srv.run(reqs, realtime=False, on_step=lambda s: peak.append((len(s.active), s.reserved)))
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
    assert max(peak2) <= 3
    joined_mid = sum(1 for r in mixed if any(o.t_admit < r.t_admit < o.t_done for o in mixed if o is not r))
    assert joined_mid > 0


def test_failures_stay_isolated():
    print("server: malformed / oversized / cancelled requests", flush=True)
    reqs = small_stream(5, seed=2)
    ref = copy.deepcopy(reqs)
    serve.ReconServer(DEV).run(ref, realtime=False)
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
    budget = int(1.5 * max(serve.estimate_request_bytes(r.shape, r.ncoils, r.npts) for r in reqs))
    huge.kspace = torch.zeros((96, huge.npts), dtype=torch.complex64)
    huge.smaps = torch.zeros((96, huge.shape, huge.shape), dtype=torch.complex64)
    assert serve.estimate_request_bytes(huge.shape, huge.ncoils, huge.npts) > budget
    srv = serve.ReconServer(DEV, hbm_budget=budget)
    allreq = reqs + [bad, wrong, huge, cancel]
    srv.run(allreq, realtime=False)
    assert "NaN" in bad.error
    assert "smaps" in wrong.error
    assert "budget" in huge.error
    assert cancel.error == "cancelled"
    for a, b in zip(reqs, ref):
        assert a.error is None
        assert rel(a.result, b.result) < 1e-6


def test_latency_stats_and_realtime_stream():
    reqs = serve.make_stream(6, rate_hz=50.0, seed=3, sizes=(64, 96), coils=(2, 4),
                             samples=(2000, 6000), niter=3)
    serve.ReconServer(DEV).run(reqs, realtime=True)
    st = serve.latency_stats(reqs)
    assert st["completed"] == 6 and st["p50"] > 0 and st["p99"] >= st["p50"]
