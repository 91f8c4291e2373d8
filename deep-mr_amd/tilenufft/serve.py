"""Continuous-batching CG-SENSE reconstruction server on one GPU.

Requests (2D non-Cartesian multicoil k-space, coil maps, trajectory, number
of CG iterations) arrive over time. The server advances every active request
by one CG-SENSE iteration per step, and all active requests share the
launches of that step:

  SENSE expansion of each request's CG direction (one launch per image size)
  -> fused apodize / pad / pruned FFT per geometry group -> ONE interpolation
  launch over every active request (ragged work queue) -> ONE gridding launch
  -> pruned inverse FFT / crop / apodize per group -> SENSE combination and
  CG update per group (three launches)

A few requests run at a time (``max_active``, default 3): on a 400 Hz stream of
10k-200k-sample, 4-32-coil requests, three active requests already fill the
device, and more of them in lockstep only lengthen every request's solve
(measured on MI355X: 3 active requests gave 1.3x the plain path's recon/s at
half its p50 latency; 32 gave the same throughput at twice its p50).
Requests are admitted at step boundaries while a slot is free and the HBM
budget covers what they hold on the device (plan and launch tables, coil
maps, k-space, CG vectors and their share of the step arenas, counted from
the library's own figures); finished requests leave and free their budget,
so new arrivals join while others are mid-solve. Plans are kept resident and
reused for repeated trajectories (``traj_key``). A request is planned as soon
as it is submitted, by a worker thread on its own HIP stream, so planning
overlaps the steps of the requests already running.

The CG recursion is deepmr's CGStep (src/deepmr/optim/cg.py:194-212) on
E^H E x = E^H y with E = NUFFT * SENSE (src/deepmr/recon/alg/linop.py:108-148),
without density compensation. The coil combination and the CG update are the
native kernels of ``tilenufft.sense``, shared with ``run_plain``: their
reductions have a fixed order that depends only on the image size, the
interpolation and gridding of a request do not depend on which other
requests share a launch, and the FFT stages transform every coil image on
its own, so a request's result is the same bits alone, in any mix, and
through the plain one-request-at-a-time path.

Nothing in a step waits for the device: finished requests record an event
(their completion time is read from it), and plans and launch tables of
finished requests are released only when the device is idle, because
freeing device memory synchronizes.
"""

__all__ = ["ReconRequest", "ReconServer", "run_plain", "estimate_request_bytes", "request_bytes",
           "latency_stats", "make_stream"]

import collections
import concurrent.futures
import ctypes
import math
import threading
import time
import weakref
from dataclasses import dataclass

import numpy as np
import torch

from . import _core, _lib, fft as _fft
from .sense import CGState, SenseGroup, START, ITERATE


@dataclass
class ReconRequest:
    """One 2D CG-SENSE reconstruction.

    coord : (nviews, nsamples, 2) float32 in [-N/2, N/2)
    kspace: (ncoils, nviews, nsamples) complex64
    smaps : (ncoils, N, N) complex64
    """

    rid: int
    coord: np.ndarray
    kspace: torch.Tensor
    smaps: torch.Tensor
    shape: int
    niter: int = 10
    arrival: float = 0.0
    traj_key: object = None
    # filled by the server
    result: torch.Tensor = None
    error: str = None
    t_admit: float = None
    t_done: float = None
    cancelled: bool = False

    @property
    def ncoils(self):
        return int(self.kspace.shape[0])

    @property
    def npts(self):
        return int(np.prod(self.coord.shape[:-1]))

    def cancel(self):
        self.cancelled = True


def _plan_model_bytes(npts, g, width=4):
    """Device bytes of a 2D plan before it exists: coordinates (8 B) and
    interpolation entries (16 B) per sample, the gridding matrix (8 B per
    tap and sample), its row offsets (4 B per cell) and the tile offsets."""
    taps = width * width
    return int(npts * (8 + 16 + 8 * taps) + 4 * g * g + (1 << 20))


def estimate_request_bytes(n, ncoils, npts, oversamp=1.25, plan_bytes=None):
    """Device bytes a request holds while active: its plan (``plan_bytes``
    from the library when the plan exists, else a model), CG state (x, r, p,
    ap), coil maps and k-space, and its rows of the step arenas (coil images,
    oversampled grids, k-space)."""
    g = math.ceil(np.float32(oversamp) * n)
    plan = _plan_model_bytes(npts, g) if plan_bytes is None else int(plan_bytes)
    state = 8 * (4 * n * n + ncoils * n * n + ncoils * npts) + 8
    arenas = 8 * (ncoils * n * n + ncoils * g * g + ncoils * npts)
    return plan + state + arenas


def request_bytes(req, plan):
    """``estimate_request_bytes`` with the plan's own device bytes (sorted
    trajectory, bins, gridding matrix) plus its share of a batch's launch
    tables: the coil-minor sample copy of the matrix gridding (records of up
    to 16 coils, each request padded to the batch's widest record) and the
    item lists."""
    tables = 8 * 16 * -(-req.ncoils // 16) * req.npts + (1 << 20)
    return estimate_request_bytes(req.shape, req.ncoils, req.npts, plan.oversamp[-1],
                                  plan_bytes=plan.interpolator.device_bytes() + tables)


class _Active:
    """Device-resident state of one admitted request."""

    def __init__(self, req, plan, device):
        self.req = req
        self.plan = plan
        self.state = CGState(req.smaps.to(device, non_blocking=True), device)
        self.y = req.kspace.to(device, non_blocking=True).to(torch.complex64).reshape(-1).contiguous()
        self.bytes = 0

    @property
    def it(self):
        return self.state.it


def _geometry(plan):
    return (tuple(plan.shape), tuple(plan.os_shape), tuple(plan.width), tuple(np.float32(plan.beta)))


class ReconServer:
    """Continuous batching of CG-SENSE requests on one HIP device."""

    def __init__(self, device=None, hbm_budget=8 << 30, max_active=3, width=4, oversamp=1.25,
                 plan_cache=64, admission="fifo", max_wait=0.08):
        if admission not in ("fifo", "sjf"):
            raise ValueError("admission is 'fifo' or 'sjf'")
        self.device = _core.compute_device(device)
        self.budget = int(hbm_budget)
        self.max_active = int(max_active)
        # admission order: arrival ("fifo"), or the least estimated work first
        # ("sjf"), where a request that has waited longer than max_wait
        # seconds goes first (oldest first), which bounds how long large
        # requests can be passed over
        self.admission = admission
        self.max_wait = float(max_wait)
        self.width = width
        self.oversamp = oversamp
        self.queue = collections.deque()
        self.active = []
        self.done = []
        self.reserved = 0
        self._plans = collections.OrderedDict()  # traj_key -> NUFFTPlan (resident)
        self._plan_cache = plan_cache
        self._plan_lock = threading.Lock()
        self._planner = concurrent.futures.ThreadPoolExecutor(max_workers=1)
        self._plan_stream = torch.cuda.Stream(device=self.device)  # default priority
        self._pending = {}  # id(req) -> future of (plan, ready event)
        self._layout = None
        self._retired = []  # plans / launch tables of finished requests, freed when idle
        self._events = []   # (request, completion event) awaiting their time stamp
        self.steps = 0
        self.stats = dict(admitted=0, rejected=0, deferred_for_memory=0, deferred_for_slots=0, layouts=0,
                          host_layout_s=0.0, host_step_s=0.0)
        self.t0 = time.perf_counter()
        self._ev0 = None

    # ---- admission -----------------------------------------------------------
    def now(self):
        return time.perf_counter() - self.t0

    def submit(self, req):
        self.queue.append(req)
        self._pending[id(req)] = self._planner.submit(self._prepare, req)

    def _prepare(self, req):
        """Validate, plan and size a request on the planning stream (worker thread)."""
        with torch.cuda.device(self.device), torch.cuda.stream(self._plan_stream):
            self._validate(req)
            need = estimate_request_bytes(req.shape, req.ncoils, req.npts, self.oversamp)
            if need > self.budget:  # hopeless even before planning
                return None, None, need
            plan = self._plan_for(req)
            need = request_bytes(req, plan)
            ev = torch.cuda.Event()
            ev.record(self._plan_stream)
        return plan, ev, need

    def close(self):
        """Stop the planning worker (pending plans are finished first) and
        release retired device memory."""
        self._planner.shutdown(wait=True)
        self._release()

    def _plan_for(self, req):
        key = req.traj_key
        with self._plan_lock:
            if key is not None and key in self._plans:
                self._plans.move_to_end(key)
                return self._plans[key]
        plan = _fft.plan_nufft(req.coord, req.shape, self.width, self.oversamp, device=self.device)
        if plan.nframes != 1 or plan.ndim != 2:
            raise ValueError("server requests are single-frame 2D")
        if key is not None:
            with self._plan_lock:
                self._plans[key] = plan
                while len(self._plans) > self._plan_cache:
                    self._retired.append(self._plans.popitem(last=False)[1])
        return plan

    def _validate(self, req):
        if req.coord.ndim < 2 or req.coord.shape[-1] != 2:
            raise ValueError("coord must be (..., 2)")
        if req.npts < 1:
            raise ValueError("the trajectory holds no samples")
        if tuple(req.smaps.shape) != (req.ncoils, req.shape, req.shape):
            raise ValueError(f"smaps shape {tuple(req.smaps.shape)} != ({req.ncoils}, {req.shape}, {req.shape})")
        if req.kspace.numel() != req.ncoils * req.npts:
            raise ValueError("k-space does not match the trajectory")
        if not np.isfinite(req.coord).all():
            raise ValueError("coordinates contain NaN or Inf")
        if req.niter < 0:
            raise ValueError("niter must be >= 0")

    def _reject(self, req, msg):
        self.stats["rejected"] += 1
        req.error = msg
        req.t_done = self.now()
        self.done.append(req)

    def _work(self, req):
        """Relative work of a request: coil images of the samples and the
        oversampled grid, per iteration, times the iterations."""
        g = math.ceil(np.float32(self.oversamp) * req.shape)
        return req.ncoils * (req.npts + 2.0 * g * g) * (req.niter + 1)

    def _pick(self):
        """Index in the queue of the request to admit next."""
        if self.admission == "fifo" or len(self.queue) == 1:
            return 0
        now = self.now()
        for i, r in enumerate(self.queue):  # cancelled or overdue requests first, oldest first
            if r.cancelled or now - r.arrival > self.max_wait:
                return i
        return min(range(len(self.queue)), key=lambda i: self._work(self.queue[i]))

    def _admit(self):
        admitted = False
        if self.queue and len(self.active) >= self.max_active:
            self.stats["deferred_for_slots"] += 1
        while self.queue and len(self.active) < self.max_active:
            i = self._pick()
            if i:  # the pick to the front, the order of the rest kept
                r = self.queue[i]
                del self.queue[i]
                self.queue.appendleft(r)
            req = self.queue[0]
            if req.cancelled:
                self.queue.popleft()
                self._pending.pop(id(req), None)
                req.error = "cancelled"
                req.t_done = self.now()
                self.done.append(req)
                continue
            fut = self._pending.get(id(req))
            if fut is None:  # not submitted through submit()
                fut = self._pending[id(req)] = self._planner.submit(self._prepare, req)
            try:
                plan, ready, need = fut.result()
            except Exception as exc:  # a malformed request fails alone
                self.queue.popleft()
                self._pending.pop(id(req), None)
                self._reject(req, f"{type(exc).__name__}: {exc}")
                continue
            if need > self.budget:
                self.queue.popleft()
                self._pending.pop(id(req), None)
                self._reject(req, f"request needs {need} bytes, more than the HBM budget {self.budget}")
                continue
            if self.reserved + need > self.budget:
                self.stats["deferred_for_memory"] += 1
                break  # wait for memory to free up (FIFO)
            self.queue.popleft()
            self._pending.pop(id(req), None)
            torch.cuda.current_stream(self.device).wait_event(ready)
            try:
                st = _Active(req, plan, self.device)
            except Exception as exc:
                self._reject(req, f"{type(exc).__name__}: {exc}")
                continue
            st.bytes = need
            self.reserved += need
            self.stats["admitted"] += 1
            req.t_admit = self.now()
            self.active.append(st)
            self._invalidate()
            admitted = True
        return admitted

    # ---- one step ------------------------------------------------------------
    def _build_layout(self):
        """Arenas, SENSE groups and batched launch tables for the active set.

        Requests with the same plan geometry (image size, oversampled grid,
        kernel) form a group: their coil images are contiguous rows of the
        group's image arena, so each step costs a fixed number of launches
        per group, and all groups share one interpolation and one gridding
        launch over the grid and k-space arenas."""
        act = sorted(self.active, key=lambda s: (s.req.shape, s.req.rid))
        dev = self.device
        groups = collections.OrderedDict()
        for s in act:
            groups.setdefault(_geometry(s.plan), []).append(s)
        grid_elems = sum(s.req.ncoils * int(np.prod(s.plan.os_shape)) for s in act)
        ksp_elems = sum(s.req.ncoils * s.req.npts for s in act)
        lay = dict(groups=[], order=act)
        lay["grid"] = torch.empty(max(grid_elems, 1), dtype=torch.complex64, device=dev)
        lay["kspace"] = torch.empty(max(ksp_elems, 1), dtype=torch.complex64, device=dev)
        goff = koff = 0
        gofs, kofs = [], []
        for key, members in groups.items():
            plan0 = members[0].plan
            n0, n1 = plan0.shape
            g0, g1 = plan0.os_shape
            rows = sum(s.req.ncoils for s in members)
            img_off, r = [], 0
            for s in members:
                img_off.append(r * n0 * n1)
                r += s.req.ncoils
            grp = dict(members=members, rows=rows, native=plan0.interpolator, os_shape=(g0, g1),
                       gslice=(goff, goff + rows * g0 * g1),
                       images=torch.empty((rows, n0, n1), dtype=torch.complex64, device=dev),
                       sense=SenseGroup([s.state for s in members], img_off, dev),
                       kslices=[])
            for s in members:
                gofs.append(goff)
                kofs.append(koff)
                grp["kslices"].append((s, koff))
                goff += s.req.ncoils * g0 * g1
                koff += s.req.ncoils * s.req.npts
            lay["groups"].append(grp)
        order = [s for grp in lay["groups"] for s in grp["members"]]
        lay["batch"] = _NativeBatch([s.plan for s in order], [s.req.ncoils for s in order], gofs, kofs, dev)
        return lay

    def step(self):
        """Admit what fits, then advance every active request by one iteration."""
        self._admit()
        for s in list(self.active):
            if s.req.cancelled:
                self._finish(s, error="cancelled")
        if not self.active:
            self._release()
            return False
        t_host = time.perf_counter()
        if self._layout is None:
            if self._ev0 is None:
                self._ev0 = torch.cuda.Event(enable_timing=True)
                self._ev0.record(torch.cuda.current_stream(self.device))
                self._t_ev0 = self.now()
            self._layout = self._build_layout()
            self.stats["layouts"] += 1
            self.stats["host_layout_s"] += time.perf_counter() - t_host
        lay = self._layout
        with torch.cuda.device(self.device):
            any_iter = False
            for grp in lay["groups"]:
                modes = [START if s.it < 0 else ITERATE for s in grp["members"]]
                grp["sense"].set_modes(modes)
                grp["any_iter"] = ITERATE in modes
                any_iter |= grp["any_iter"]
            if any_iter:
                # forward NUFFT of the CG directions: SENSE expansion, pad +
                # FFT per group, one interpolation launch
                for grp in lay["groups"]:
                    if not grp["any_iter"]:
                        continue
                    grp["sense"].expand(grp["images"])
                    a, b = grp["gslice"]
                    _fft_forward(grp["native"], grp["images"], lay["grid"][a:b].view(-1, *grp["os_shape"]),
                                 grp["rows"])
                lay["batch"].interp(lay["grid"], lay["kspace"])
            # measured k-space of requests starting this step (their
            # interpolation output, if any, is overwritten here)
            for grp in lay["groups"]:
                for s, ko in grp["kslices"]:
                    if s.it < 0:
                        lay["kspace"][ko:ko + s.y.numel()].copy_(s.y)
            # adjoint of every request: one gridding launch, IFFT + crop per
            # group, then coil combination and the CG update
            lay["batch"].grid(lay["kspace"], lay["grid"])
            for grp in lay["groups"]:
                a, b = grp["gslice"]
                _fft_adjoint(grp["native"], lay["grid"][a:b].view(-1, *grp["os_shape"]), grp["images"],
                             grp["rows"])
                grp["sense"].combine(grp["images"])
                grp["sense"].cg_step()
            finished = []
            for s in self.active:
                s.state.it += 1
                if s.state.it >= s.req.niter:
                    finished.append(s)
            for s in finished:
                self._finish(s)
        self.steps += 1
        self.stats["host_step_s"] += time.perf_counter() - t_host
        return True

    def _finish(self, s, error=None):
        self.active.remove(s)
        self.reserved -= s.bytes
        req = s.req
        req.error = error
        if error is None:
            req.result = s.state.x.clone()
            ev = torch.cuda.Event(enable_timing=True)
            ev.record(torch.cuda.current_stream(self.device))
            self._events.append((req, ev))
        else:
            req.t_done = self.now()
        if req.traj_key is None:
            self._retired.append(s.plan)
        s.state = None
        self.done.append(req)
        self._invalidate()

    def _invalidate(self):
        """Drop the layout of the active set; its launch tables are freed
        later (freeing device memory synchronizes)."""
        if self._layout is not None:
            self._retired.append(self._layout)
            self._layout = None

    def _release(self):
        """Time-stamp finished requests and free retired plans and launch
        tables (the device is idle or about to be synchronized anyway)."""
        torch.cuda.synchronize(self.device)
        for req, ev in self._events:
            req.t_done = self._t_ev0 + self._ev0.elapsed_time(ev) * 1e-3 if self._ev0 is not None else self.now()
        self._events.clear()
        self._retired.clear()

    # ---- driving a stream ----------------------------------------------------
    def run(self, requests, realtime=True, on_step=None):
        """Serve ``requests`` (arrival times in seconds from now). Returns them
        with ``result`` / ``error`` and timing fields filled."""
        pending = sorted(requests, key=lambda r: r.arrival)
        self.t0 = time.perf_counter()
        self._ev0 = None
        i = 0
        total = len(pending)
        while len(self.done) < total:
            now = self.now()
            while i < total and (not realtime or pending[i].arrival <= now):
                self.submit(pending[i])
                i += 1
            progressed = self.step()
            if on_step is not None:
                on_step(self)
            if not progressed and i < total and realtime:
                wait = pending[i].arrival - self.now()
                if wait > 0:
                    time.sleep(min(wait, 0.05))
            elif not progressed and not self.queue and i >= total:
                break
        self._release()
        return requests


def _fft_forward(native, images, grid, nimg):
    if native.has_fft:
        native.fft_forward(images, grid, nimg)
    else:
        native.apod_pad(images, grid, nimg)
        grid.copy_(torch.fft.fftn(grid, dim=(-2, -1)))


def _fft_adjoint(native, grid, images, nimg):
    if native.has_fft:
        native.fft_adjoint(grid, images, nimg)
    else:
        native.crop_apod(torch.fft.ifftn(grid, dim=(-2, -1)), images, nimg)


class _NativeBatch:
    """tn_batch over arbitrary plans with explicit arena offsets."""

    def __init__(self, plans, ncoils, grid_off, samp_off, device):
        n = len(plans)
        lib = _lib.load()
        handles = (ctypes.c_void_p * n)(*[p.interpolator.handle.value for p in plans])
        nb = (ctypes.c_int64 * n)(*ncoils)
        go = (ctypes.c_int64 * n)(*grid_off)
        so = (ctypes.c_int64 * n)(*samp_off)
        out = ctypes.c_void_p()
        with torch.cuda.device(device):
            _lib.check(lib.tn_batch_create(handles, nb, go, so, None, n, ctypes.byref(out)), "tn_batch_create")
        self.handle = out
        self.device = device
        self.plans = plans
        from .batch import _destroy_batch

        self._fin = weakref.finalize(self, _destroy_batch, out.value, device)

    def interp(self, grid, kspace):
        _lib.check(_lib.load().tn_batch_interp(self.handle, _core.ptr(grid), _core.ptr(kspace), None,
                                               _core.stream_ptr(self.device)), "tn_batch_interp")

    def grid(self, kspace, grid):
        _lib.check(_lib.load().tn_batch_grid(self.handle, _core.ptr(kspace), _core.ptr(grid), None,
                                             _core.stream_ptr(self.device)), "tn_batch_grid")


def cg_sense_plain(req, plan, device, replan=False, width=4, oversamp=1.25):
    """One request's CG-SENSE through the plain NUFFT (apply_nufft /
    apply_nufft_adj with the request's plan, or the per-call planning
    nufft / nufft_adj when ``replan``) and the same native SENSE / CG
    kernels as the server. Returns the image (on the device)."""
    dev = device
    state = CGState(req.smaps.to(dev), dev)
    group = SenseGroup([state], [0], dev)
    y = req.kspace.to(dev, torch.complex64).reshape(req.ncoils, *req.coord.shape[:-1])

    def fwd(v):
        if replan:
            return _fft.nufft(v, req.coord, req.shape, device=dev, width=width, oversamp=oversamp)
        return _fft.apply_nufft(v, plan)

    def adj(k):
        if replan:
            return _fft.nufft_adj(k, req.coord, req.shape, device=dev, width=width, oversamp=oversamp)
        return _fft.apply_nufft_adj(k, plan)

    img = adj(y).contiguous()
    group.set_modes([START])
    group.combine(img)
    group.cg_step()
    for _ in range(req.niter):
        group.set_modes([ITERATE])
        group.expand(img)
        img = adj(fwd(img)).contiguous()
        group.combine(img)
        group.cg_step()
    return state.x.clone()


def run_plain(requests, device=None, width=4, oversamp=1.25, realtime=True, replan=False):
    """The same requests one at a time through the plain GPU NUFFT.

    Each request runs its CG-SENSE with apply_nufft / apply_nufft_adj calls
    (one request on the device at a time, plan reused across its iterations
    unless ``replan``, where every call plans like deepmr's nufft/nufft_adj),
    with the server's SENSE / CG kernels. Completion times come from events
    on the stream, as in the server.
    """
    dev = _core.compute_device(device)
    t0 = time.perf_counter()
    stream = torch.cuda.current_stream(dev)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    t_ev0 = time.perf_counter() - t0
    events = []
    plans = []
    for req in sorted(requests, key=lambda r: r.arrival):
        if realtime:
            wait = req.arrival - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
        req.t_admit = time.perf_counter() - t0
        try:
            with torch.cuda.device(dev):
                plan = None if replan else _fft.plan_nufft(req.coord, req.shape, width, oversamp, device=dev)
                req.result = cg_sense_plain(req, plan, dev, replan, width, oversamp)
                ev = torch.cuda.Event(enable_timing=True)
                ev.record(stream)
                events.append((req, ev))
                plans.append(plan)  # freed after the stream has drained (freeing synchronizes)
        except Exception as exc:
            req.error = f"{type(exc).__name__}: {exc}"
            req.t_done = time.perf_counter() - t0
    torch.cuda.synchronize(dev)
    for req, ev in events:
        req.t_done = t_ev0 + ev0.elapsed_time(ev) * 1e-3
    plans.clear()
    return requests


def latency_stats(requests):
    """p50 / p99 latency (s) and completed recon/s over the stream."""
    ok = [r for r in requests if r.error is None and r.t_done is not None]
    if not ok:
        return dict(p50=None, p99=None, recon_per_s=0.0, completed=0)
    lat = np.array([r.t_done - r.arrival for r in ok])
    span = max(r.t_done for r in ok) - min(r.arrival for r in requests)
    return dict(p50=float(np.percentile(lat, 50)), p99=float(np.percentile(lat, 99)),
                recon_per_s=len(ok) / span if span > 0 else float("inf"), completed=len(ok))


def make_stream(nreq, rate_hz, seed=0, sizes=(128, 192, 256, 320), coils=(4, 32),
                samples=(10_000, 200_000), niter=10, device="cpu", traj_pool=None):
    """Synthetic request stream: Poisson arrivals at ``rate_hz``; per request a
    matrix size from ``sizes``, U{coils} coils, a radial or spiral trajectory
    with U[samples] samples per coil, complex N(0,1) k-space and Gaussian-bump
    coil maps. ``traj_pool`` > 0 draws trajectories from that many shared
    keys per size (repeated trajectories reuse resident plans)."""
    from . import testing

    rng = np.random.default_rng(seed)
    gen = torch.Generator().manual_seed(seed)
    t = 0.0
    out = []
    maps_cache = {}
    for i in range(nreq):
        t += rng.exponential(1.0 / rate_hz) if rate_hz > 0 else 0.0
        n = int(rng.choice(sizes))
        c = int(rng.integers(coils[0], coils[1] + 1))
        m = int(rng.integers(samples[0], samples[1] + 1))
        kind = "radial" if rng.random() < 0.5 else "spiral"
        key = None
        trng = rng
        if traj_pool:
            j = int(rng.integers(traj_pool))
            key = (kind, n, j)
            trng = np.random.default_rng([seed, 0 if kind == "radial" else 1, n, j])
            m = int(trng.integers(samples[0], samples[1] + 1))
        if kind == "radial":
            ns = 2 * n
            nv = max(1, m // ns)
            coord = testing.golden_angle_radial(nv, ns, n, offset=float(trng.uniform(0, np.pi)))
        else:
            ns = 1024
            nv = max(1, m // ns)
            coord = testing.vd_spiral(nv, ns, n, rotation=float(trng.uniform(0, 2 * np.pi)))
        pin = torch.device(device).type == "cpu" and torch.cuda.is_available()
        if (c, n) not in maps_cache:
            mp = torch.from_numpy(testing.gaussian_coil_maps(c, (n, n), seed=c))
            maps_cache[(c, n)] = mp.pin_memory() if pin else mp.to(device)
        k = testing.complex_normal((c, *coord.shape[:-1]), generator=gen)
        k = k.pin_memory() if pin else k.to(device)  # pinned host data copies asynchronously
        out.append(ReconRequest(rid=i, coord=coord, kspace=k, smaps=maps_cache[(c, n)],
                                shape=n, niter=niter, arrival=t, traj_key=key))
    return out
