"""Admission order of the CG-SENSE server (CPU only, no device): first in,
first out by default; least estimated work first with ``sjf``, except that
cancelled requests and requests waiting longer than ``max_wait`` go first,
oldest first."""

import collections
import types

import numpy as np

from tilenufft import serve


def _req(rid, ncoils, npts, n=128, arrival=0.0):
    return serve.ReconRequest(rid=rid, coord=np.zeros((1, npts, 2), np.float32),
                              kspace=np.zeros((ncoils, 1, npts), np.complex64),
                              smaps=None, shape=n, arrival=arrival)


def _server(admission, queue, now, max_wait=0.08):
    srv = types.SimpleNamespace(admission=admission, max_wait=max_wait, oversamp=1.25,
                                queue=collections.deque(queue), now=lambda: now)
    srv._work = types.MethodType(serve.ReconServer._work, srv)
    return srv


def test_fifo_takes_the_head():
    q = [_req(0, 32, 200_000), _req(1, 4, 10_000)]
    assert serve.ReconServer._pick(_server("fifo", q, now=1.0)) == 0


def test_sjf_takes_the_least_work_unless_one_is_overdue():
    q = [_req(0, 32, 200_000, arrival=0.10), _req(1, 4, 10_000, arrival=0.11), _req(2, 8, 10_000, arrival=0.12)]
    assert serve.ReconServer._pick(_server("sjf", q, now=0.15)) == 1
    # the head has waited past max_wait: it goes first
    assert serve.ReconServer._pick(_server("sjf", q, now=0.19)) == 0
    q[2].cancel()
    assert serve.ReconServer._pick(_server("sjf", q, now=0.15)) == 2


def test_work_counts_coils_samples_and_grid():
    srv = _server("sjf", [], now=0.0)
    a, b = _req(0, 4, 10_000), _req(1, 8, 10_000)
    assert srv._work(b) == 2 * srv._work(a)
    assert srv._work(_req(2, 4, 10_000, n=256)) > srv._work(a)
