"""Continuous-batching CG-SENSE reconstruction server on one GPU.

Requests (2D non-Cartesian multicoil k-space, coil maps, trajectory, number
of CG iterations) arrive over time. The server advances every active request
by one CG-SENSE iteration per step, and all active requests share the
launches of that step:

  coil expansion -> per-geometry apod+pad -> per-geometry batched rocFFT ->
  ONE interpolation launch over every active request (ragged work queue) ->
  ONE gridding launch over every active request -> per-geometry batched
  inverse rocFFT -> per-geometry crop+apod -> coil combination -> CG update

Requests are admitted at step boundaries while a slot is free and the HBM
budget covers their resident state (plan / sorted trajectory, coil maps,
k-space, CG vectors and their share of the step arenas); finished requests
leave and free their budget, so new arrivals join while others are mid-solve.
Plans are kept resident and reused for repeated trajectories (``traj_key``).
A request is planned as soon as it is submitted, by a worker thread on its
own HIP stream, so planning (host arithmetic, binning and sorting on the
device) overlaps the steps of the requests already running; its k-space and
coil maps are copied to the device without blocking at admission.

The CG recursion is deepmr's CGStep (src/deepmr/optim/cg.py:194-212) on
E^H E x = E^H y with E = NUFFT * SENSE (src/deepmr/recon/alg/linop.py:108-148),
without density compensation.

The per-request interpolation and gridding results do not depend on which
other requests share a launch (each work item is one request's sample block
or tile); only the batched FFT may pick another algorithm, so a result alone
and in a mix agree to float32 round-off.
"""

__all__ = ["ReconRequest", "ReconServer", "run_plain", "estimate_request_bytes", "latency_stats",
           "make_stream"]

import collections
import concurrent.futures
import math
import threading
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _core, fft as _fft
from .batch import BatchedNUFFT


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


def estimate_request_bytes(n, ncoils, npts, oversamp=1.25, dup=1.35):
    """Device bytes a request holds while active."""
    g = math.ceil(np.float32(oversamp) * n)
    plan = 8 * npts + int(16 * dup * npts) + 4 * 512
    state = 8 * (ncoils * n * n  # coil maps
                 + ncoils * npts  # measured k-space
                 + 5 * n * n)  # x, r, p, AHAp, b
    arenas = 8 * (ncoils * n * n + 2 * ncoils * g * g + ncoils * npts)
    return plan + state + arenas


class _Active:
    """Device-resident state of one admitted request."""

    def __init__(self, req, plan, device):
        self.req = req
        self.plan = plan
        # pinned host tensors copy without blocking the host
        self.smaps = req.smaps.to(device, non_blocking=True).to(torch.complex64).contiguous()
        self.y = req.kspace.to(device, non_blocking=True).to(torch.complex64).reshape(req.ncoils, -1).contiguous()
        n = req.shape
        self.x = torch.zeros((n, n), dtype=torch.complex64, device=device)
        self.r = None
        self.p = None
        self.rs_old = None
        self.it = -1  # -1: E^H y pending, then 0..niter-1
        self.bytes = 0


class ReconServer:
    """Continuous batching of CG-SENSE requests on one HIP device."""

    def __init__(self, device=None, hbm_budget=8 << 30, max_active=32, width=4, oversamp=1.25,
                 plan_cache=64):
        self.device = _core.compute_device(device)
        self.budget = int(hbm_budget)
        self.max_active = int(max_active)
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
        self._layout_key = None
        self._layout = None
        self.steps = 0
        self.stats = dict(admitted=0, rejected=0, deferred_for_memory=0, deferred_for_slots=0)
        self.t0 = time.perf_counter()

    # ---- admission -----------------------------------------------------------
    def now(self):
        return time.perf_counter() - self.t0

    def submit(self, req):
        self.queue.append(req)
        self._pending[id(req)] = self._planner.submit(self._prepare, req)

    def _prepare(self, req):
        """Validate and plan a request on the planning stream (worker thread)."""
        with torch.cuda.device(self.device), torch.cuda.stream(self._plan_stream):
            self._validate(req)
            plan = self._plan_for(req)
            ev = torch.cuda.Event()
            ev.record(self._plan_stream)
        return plan, ev

    def close(self):
        """Stop the planning worker (pending plans are finished first)."""
        self._planner.shutdown(wait=True)

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
                    self._plans.popitem(last=False)
        return plan

    def _validate(self, req):
        if req.coord.ndim < 2 or req.coord.shape[-1] != 2:
            raise ValueError("coord must be (..., 2)")
        if tuple(req.smaps.shape) != (req.ncoils, req.shape, req.shape):
            raise ValueError(f"smaps shape {tuple(req.smaps.shape)} != ({req.ncoils}, {req.shape}, {req.shape})")
        if req.kspace.numel() != req.ncoils * req.npts:
            raise ValueError("k-space does not match the trajectory")
        if not np.isfinite(req.coord).all():
            raise ValueError("coordinates contain NaN or Inf")

    def _admit(self):
        admitted = False
        if self.queue and len(self.active) >= self.max_active:
            self.stats["deferred_for_slots"] += 1
        while self.queue and len(self.active) < self.max_active:
            req = self.queue[0]
            if req.cancelled:
                self.queue.popleft()
                self._pending.pop(id(req), None)
                req.error = "cancelled"
                req.t_done = self.now()
                self.done.append(req)
                continue
            need = estimate_request_bytes(req.shape, req.ncoils, req.npts, self.oversamp)
            if need > self.budget:
                self.queue.popleft()
                self._pending.pop(id(req), None)
                req.error = f"request needs {need} bytes, more than the HBM budget {self.budget}"
                self.stats["rejected"] += 1
                req.t_done = self.now()
                self.done.append(req)
                continue
            if self.reserved + need > self.budget:
                self.stats["deferred_for_memory"] += 1
                break  # wait for memory to free up (FIFO)
            self.queue.popleft()
            try:
                fut = self._pending.pop(id(req), None)
                if fut is None:  # not submitted through submit()
                    fut = self._planner.submit(self._prepare, req)
                plan, ready = fut.result()
                torch.cuda.current_stream(self.device).wait_event(ready)
                st = _Active(req, plan, self.device)
            except Exception as exc:  # a malformed request fails alone
                self.stats["rejected"] += 1
                req.error = f"{type(exc).__name__}: {exc}"
                req.t_done = self.now()
                self.done.append(req)
                continue
            st.bytes = need
            self.reserved += need
            self.stats["admitted"] += 1
            req.t_admit = self.now()
            self.active.append(st)
            admitted = True
        return admitted

    # ---- one step ------------------------------------------------------------
    def _build_layout(self):
        """Arenas, stacked CG state and batched launch tables for the active set.

        Requests with the same matrix size form a group: their coil images,
        padded grids and CG vectors are stacked so that each step costs a
        fixed number of launches per group, and all groups share one
        interpolation and one gridding launch.
        """
        self._store_state()
        act = sorted(self.active, key=lambda s: (s.req.shape, s.req.rid))
        dev = self.device
        groups = collections.OrderedDict()
        for s in act:
            groups.setdefault(s.req.shape, []).append(s)
        grid_elems = sum(s.req.ncoils * math.ceil(np.float32(self.oversamp) * s.req.shape) ** 2 for s in act)
        ksp_elems = sum(s.req.ncoils * s.req.npts for s in act)
        lay = dict(order=act, groups=[])
        lay["gridA"] = torch.empty(grid_elems, dtype=torch.complex64, device=dev)
        lay["gridB"] = torch.empty(grid_elems, dtype=torch.complex64, device=dev)
        lay["kspace"] = torch.empty(max(ksp_elems, 1), dtype=torch.complex64, device=dev)
        goff = koff = 0
        offs = {}
        for n, members in groups.items():
            g = math.ceil(np.float32(self.oversamp) * n)
            rows = sum(s.req.ncoils for s in members)
            row2req = torch.cat([torch.full((s.req.ncoils,), i, dtype=torch.long)
                                 for i, s in enumerate(members)]).to(dev)
            init = torch.tensor([s.it < 0 for s in members], device=dev)
            grp = dict(n=n, g=g, members=members, rows=rows, row2req=row2req, init=init,
                       any_init=bool(init.any()), any_cg=bool((~init).any()),
                       gslice=(goff, goff + rows * g * g),
                       smaps=torch.cat([s.smaps for s in members]),
                       images=torch.empty((rows, n, n), dtype=torch.complex64, device=dev))
            grp["X"] = torch.stack([s.x for s in members])
            grp["R"] = torch.stack([s.r if s.r is not None else s.x for s in members])
            grp["P"] = torch.stack([s.p if s.p is not None else s.x for s in members])
            grp["rs"] = torch.stack([s.rs_old if s.rs_old is not None else
                                     torch.ones((), device=dev) for s in members]).reshape(-1)
            for s in members:
                offs[s.req.rid] = (goff, koff)
                goff += s.req.ncoils * g * g
                koff += s.req.ncoils * s.req.npts
            lay["groups"].append(grp)
        lay["offs"] = offs
        lay["grid_batch"] = _NativeBatch([s.plan for s in act], [s.req.ncoils for s in act],
                                         [offs[s.req.rid][0] for s in act],
                                         [offs[s.req.rid][1] for s in act], dev)
        cg = [s for s in act if s.it >= 0]
        lay["interp_batch"] = (_NativeBatch([s.plan for s in cg], [s.req.ncoils for s in cg],
                                            [offs[s.req.rid][0] for s in cg],
                                            [offs[s.req.rid][1] for s in cg], dev) if cg else None)
        return lay

    def _store_state(self):
        """Write the stacked CG state of the current layout back to the requests."""
        lay = self._layout
        if lay is None:
            return
        for grp in lay["groups"]:
            for i, s in enumerate(grp["members"]):
                if s.it < 0 or s not in self.active:
                    continue
                s.x, s.r, s.p = grp["X"][i].clone(), grp["R"][i].clone(), grp["P"][i].clone()
                s.rs_old = grp["rs"][i].clone()
        self._layout = None
        self._layout_key = None

    def _layout_for_step(self):
        key = tuple((s.req.rid, s.it >= 0) for s in sorted(self.active, key=lambda s: (s.req.shape, s.req.rid)))
        if key != self._layout_key:
            self._layout = self._build_layout()
            self._layout_key = key
        return self._layout

    def step(self):
        """Admit what fits, then advance every active request by one iteration."""
        self._admit()
        for s in list(self.active):
            if s.req.cancelled:
                self._finish(s, error="cancelled")
        if not self.active:
            return False
        lay = self._layout_for_step()
        with torch.cuda.device(self.device):
            # forward NUFFT of the search directions (CG requests): coil
            # expansion, pad + FFT per group, one interpolation launch
            if lay["interp_batch"] is not None:
                for grp in lay["groups"]:
                    a, b = grp["gslice"]
                    g = grp["g"]
                    if grp["any_cg"]:
                        torch.mul(grp["smaps"], grp["P"].index_select(0, grp["row2req"]), out=grp["images"])
                    nat = grp["members"][0].plan.interpolator
                    if nat.has_fft:
                        nat.fft_forward(grp["images"], lay["gridB"][a:b].view(-1, g, g), grp["rows"])
                    else:
                        ga = lay["gridA"][a:b].view(-1, g, g)
                        nat.apod_pad(grp["images"], ga, grp["rows"])
                        torch.fft.fftn(ga, dim=(-2, -1), out=lay["gridB"][a:b].view(-1, g, g))
                lay["interp_batch"].interp(lay["gridB"], lay["kspace"])
            # measured k-space of requests starting this step
            for grp in lay["groups"]:
                if grp["any_init"]:
                    for s in grp["members"]:
                        if s.it < 0:
                            _, ko = lay["offs"][s.req.rid]
                            lay["kspace"][ko:ko + s.y.numel()].copy_(s.y.reshape(-1))
            # adjoint of every request: one gridding launch, IFFT + crop per group
            lay["grid_batch"].grid(lay["kspace"], lay["gridA"])
            for grp in lay["groups"]:
                a, b = grp["gslice"]
                g = grp["g"]
                nat = grp["members"][0].plan.interpolator
                if nat.has_fft:  # gridA is scratch from here on
                    nat.fft_adjoint(lay["gridA"][a:b].view(-1, g, g), grp["images"], grp["rows"])
                else:
                    gb = lay["gridB"][a:b].view(-1, g, g)
                    torch.fft.ifftn(lay["gridA"][a:b].view(-1, g, g), dim=(-2, -1), out=gb)
                    nat.crop_apod(gb, grp["images"], grp["rows"])
                self._cg_update(grp)
            finished = []
            for grp in lay["groups"]:
                for s in grp["members"]:
                    s.it += 1
                    if s.it >= s.req.niter:
                        finished.append(s)
            if finished:
                torch.cuda.synchronize(self.device)
                self._store_state()
                for s in finished:
                    self._finish(s)
        self.steps += 1
        return True

    @staticmethod
    def _cg_update(grp):
        """Coil combination and one CGStep (src/deepmr/optim/cg.py:203-212) for
        every request of the group; requests starting this step set
        r = p = E^H y instead."""
        nreq = len(grp["members"])
        n = grp["n"]
        ap = torch.zeros((nreq, n, n), dtype=torch.complex64, device=grp["images"].device)
        ap.index_add_(0, grp["row2req"], grp["smaps"].conj() * grp["images"])
        X, R, P, rs = grp["X"], grp["R"], grp["P"], grp["rs"]
        if grp["any_cg"]:
            alpha = rs / (P.conj() * ap).sum(dim=(-2, -1))
            a3 = alpha[:, None, None]
            x_new = X + P * a3
            r_new = R + ap * (-a3)
            rs_new = torch.real((r_new.conj() * r_new).sum(dim=(-2, -1)))
            p_new = r_new + P * (rs_new / rs)[:, None, None]
        if grp["any_init"]:
            rs_init = torch.real((ap.conj() * ap).sum(dim=(-2, -1)))
            if grp["any_cg"]:
                m = grp["init"]
                m3 = m[:, None, None]
                x_new = torch.where(m3, torch.zeros_like(X), x_new)
                r_new = torch.where(m3, ap, r_new)
                p_new = torch.where(m3, ap, p_new)
                rs_new = torch.where(m, rs_init, rs_new)
            else:
                x_new, r_new, p_new, rs_new = torch.zeros_like(X), ap, ap.clone(), rs_init
        X.copy_(x_new)
        R.copy_(r_new)
        P.copy_(p_new)
        rs.copy_(rs_new)

    def _finish(self, s, error=None):
        if self._layout is not None:
            self._store_state()
        self.active.remove(s)
        self.reserved -= s.bytes
        s.req.result = None if error else s.x
        s.req.error = error
        s.req.t_done = self.now()
        self.done.append(s.req)

    # ---- driving a stream ----------------------------------------------------
    def run(self, requests, realtime=True, on_step=None):
        """Serve ``requests`` (arrival times in seconds from now). Returns them
        with ``result`` / ``error`` and timing fields filled."""
        pending = sorted(requests, key=lambda r: r.arrival)
        self.t0 = time.perf_counter()
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
        torch.cuda.synchronize(self.device)
        return requests


class _NativeBatch:
    """tn_batch over arbitrary plans with explicit arena offsets."""

    def __init__(self, plans, ncoils, grid_off, samp_off, device):
        import ctypes
        import weakref

        from . import _lib

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
        from . import _lib

        _lib.check(_lib.load().tn_batch_interp(self.handle, _core.ptr(grid), _core.ptr(kspace), None,
                                               _core.stream_ptr(self.device)), "tn_batch_interp")

    def grid(self, kspace, grid):
        from . import _lib

        _lib.check(_lib.load().tn_batch_grid(self.handle, _core.ptr(kspace), _core.ptr(grid), None,
                                             _core.stream_ptr(self.device)), "tn_batch_grid")


def run_plain(requests, device=None, width=4, oversamp=1.25, realtime=True, replan=False):
    """The same requests one at a time through the plain GPU NUFFT.

    Each request runs its CG-SENSE with apply_nufft / apply_nufft_adj calls
    (one request on the device at a time, plan reused across its iterations
    unless ``replan``, where every call plans like deepmr's nufft/nufft_adj).
    """
    dev = _core.compute_device(device)
    t0 = time.perf_counter()
    for req in sorted(requests, key=lambda r: r.arrival):
        if realtime:
            wait = req.arrival - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
        req.t_admit = time.perf_counter() - t0
        try:
            with torch.cuda.device(dev):
                smaps = req.smaps.to(dev, torch.complex64)
                y = req.kspace.to(dev, torch.complex64).reshape(req.ncoils, *req.coord.shape[:-1])
                plan = None if replan else _fft.plan_nufft(req.coord, req.shape, width, oversamp, device=dev)

                def fwd(v):
                    if replan:
                        return _fft.nufft(v, req.coord, req.shape, device=dev, width=width, oversamp=oversamp)
                    return _fft.apply_nufft(v, plan)

                def adj(k):
                    if replan:
                        return _fft.nufft_adj(k, req.coord, req.shape, device=dev, width=width,
                                              oversamp=oversamp)
                    return _fft.apply_nufft_adj(k, plan)

                def EHE(v):
                    return (smaps.conj() * adj(fwd(smaps * v))).sum(0)

                b = (smaps.conj() * adj(y)).sum(0)
                x = torch.zeros_like(b)
                r = b
                p = b.clone()
                rs_old = torch.real(torch.vdot(r.reshape(-1), r.reshape(-1)))
                for _ in range(req.niter):
                    ap = EHE(p)
                    alpha = rs_old / torch.vdot(p.reshape(-1), ap.reshape(-1))
                    x = x + p * alpha
                    r = r + ap * (-alpha)
                    rs_new = torch.real(torch.vdot(r.reshape(-1), r.reshape(-1)))
                    p = r + p * (rs_new / rs_old)
                    rs_old = rs_new
                torch.cuda.synchronize(dev)
            req.result = x
        except Exception as exc:
            req.error = f"{type(exc).__name__}: {exc}"
        req.t_done = time.perf_counter() - t0
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
            m = torch.from_numpy(testing.gaussian_coil_maps(c, (n, n), seed=c))
            maps_cache[(c, n)] = m.pin_memory() if pin else m.to(device)
        k = testing.complex_normal((c, *coord.shape[:-1]), generator=gen)
        k = k.pin_memory() if pin else k.to(device)  # pinned host data copies asynchronously
        out.append(ReconRequest(rid=i, coord=coord, kspace=k, smaps=maps_cache[(c, n)],
                                shape=n, niter=niter, arrival=t, traj_key=key))
    return out
