"""NUFFT benchmark on one MI355X.

Workload (default): batched 2D multicoil spiral — 8 frames x 16 coils, 256x256
complex64 images, one variable-density spiral per frame with 24..48
interleaves x 2048 samples (uneven sample counts, golden-angle rotation per
frame), Gaussian-bump coil maps, oversampling 2.0, Kaiser-Bessel width 4.
Tile / chunk sizes come from the autotuner; the forward+adjoint pair of all
frames and coils is one HIP graph. One step = forward + adjoint of every
sample of every frame and coil, inputs resident in HBM.

Prints one JSON line with samples/s, the roofline of the dominant kernel
(interpolation or gridding; achieved = algorithmic bytes / average kernel
time from HIP events on the launch stream) and a CPU baseline (the oracle's
C/OpenMP restatement of deepmr's Numba CPU loops) on a bounded sample.
"""

import argparse
import gc
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "deep-mr_amd"), ROOT):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402
import torch  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "NUFFT samples/s fwd+adj, % HBM roofline; served recon p50/p99 vs plain"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_spiral_workload(device, nframes=8, ncoils=16, n=256, nsamp=2048, osf=2.0, width=4, seed=0):
    import tilenufft
    from tilenufft import testing

    rng = np.random.default_rng(seed)
    inter = rng.integers(24, 49, size=nframes)
    coords = [testing.vd_spiral(int(inter[f]), nsamp, n, rotation=f * testing.GOLDEN_ANGLE)
              for f in range(nframes)]
    plans = [tilenufft.fft.plan_nufft(c, n, width, osf, device=device) for c in coords]
    smaps = torch.from_numpy(testing.gaussian_coil_maps(ncoils, (n, n), seed=seed)).to(device)
    gen = torch.Generator().manual_seed(seed)
    frames = testing.complex_normal((nframes, n, n), generator=gen).to(device)
    return dict(coords=coords, plans=plans, smaps=smaps, frames=frames, interleaves=inter.tolist(),
                n=n, ncoils=ncoils, osf=osf, width=width, nsamp=nsamp)


def variant_key(plan, kind, ncoils):
    """Launch variant of a workload's dominant kernel, as the traffic file
    keys it: the effective launch configuration (requested 0 = automatic
    values resolved as the library resolves them), so equal launches get
    equal keys."""
    if kind == "grid" and plan.grid_mode in (0, 3) and plan.interpolator.grid_mode_effective() == 3:
        mc = min(plan.grid_chunk if plan.grid_chunk > 0 else 16, 16, ncoils)
        record = 2 if mc <= 2 else 4 if mc <= 4 else 8 if mc <= 8 else 16
        split = plan.grid_split
        auto_long = 256 if plan.ndim == 3 else 128  # the library's defaults
        long_row = auto_long if split == 0 else 1024 if split < 0 else min(split, 1024)
        return f"grid|matrix|record={record}|long={long_row}"
    if kind == "grid":
        return (f"grid|mode={plan.grid_mode}|tile={list(plan.interpolator.tile())}|chunk={plan.grid_chunk}"
                f"|split={plan.grid_split}")
    if kind == "interp":
        return f"interp|mode={plan.interp_mode}|itile={plan.itile}|chunk={plan.interp_chunk}"
    return kind


# kernels of each timed stage (short names as fold_counters derives them)
STAGE_KERNELS = {
    "grid": ("samp_transpose_kernel", "grid_matrix_t_kernel", "grid_matrix_kernel", "matrix_fold_kernel",
             "grid_tiles_kernel", "grid_runs_kernel", "grid_kernel", "fold_kernel"),
    "interp": ("interp_staged_kernel", "interp_kernel"),
    "fft_fwd": ("fft_cols_fwd_kernel", "fft_rows_kernel|false"),
    "fft_adj": ("fft_rows_kernel|true", "fft_cols_inv_kernel"),
}


def short_kernel_name(name):
    """`void ns::(anonymous namespace)::k<1, 2>(args)` -> `k<1, 2>`."""
    n = name.strip().replace("(anonymous namespace)::", "")
    if n.startswith("void "):
        n = n[5:]
    depth = 0
    for i, ch in enumerate(n):
        depth += ch == "<"
        depth -= ch == ">"
        if ch == "(" and depth == 0:
            n = n[:i]
            break
    base, _, targs = n.partition("<")
    return base.split("::")[-1] + ("<" + targs if targs else "")


def _stage_of(kname):
    base = kname.split("<")[0]
    for stage, names in STAGE_KERNELS.items():
        for entry in names:
            nm, _, flag = entry.partition("|")
            if base == nm and (not flag or kname.rstrip(">").endswith(flag)):
                return stage
    return None


def fold_counters(dirs, bench_line, copy_mib):
    """Per-kernel means of rocprofv3 counter runs of ``bench.py --kernels-only
    --calibrate-copy MIB`` (one directory per run), folded into the HBM bytes
    per launch of each stage. FETCH_SIZE / WRITE_SIZE are KiB per dispatch;
    each is scaled by what it reads for the plain copy of ``copy_mib`` MiB
    (the three largest-writing elementwise dispatches), so the bytes are the
    copy-calibrated estimate. ``bench_line`` is the JSON line of one of those
    runs: it names the launch variants the traffic is keyed by."""
    import csv
    import glob

    per = {}  # (kernel, counter) -> [values per dispatch]
    for d in dirs:
        files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        if not files:
            raise SystemExit(f"no counter_collection.csv under {d}")
        for fn in files:
            acc = {}
            with open(fn) as f:
                for row in csv.DictReader(f):
                    key = (row["Dispatch_Id"], short_kernel_name(row["Kernel_Name"]), row["Counter_Name"])
                    acc[key] = acc.get(key, 0.0) + float(row["Counter_Value"])
            for (_, k, c), v in acc.items():
                per.setdefault((k, c), []).append(v)
    copy_bytes = copy_mib * (1 << 20)
    scale = {}
    for c in ("FETCH_SIZE", "WRITE_SIZE"):
        cal = sorted(v for (k, cc), vs in per.items() if cc == c and "elementwise" in k for v in vs)[-3:]
        if cal:
            scale[c] = copy_bytes / (float(np.mean(cal)) * 1024.0)
    with open(bench_line) as f:
        line = json.loads([ln for ln in f if ln.startswith("{")][-1])
    variants = line.get("kernels", {}).get("variants", {})
    stages = {}
    for (k, c), vs in sorted(per.items()):
        st = _stage_of(k)
        if st is None:
            continue
        e = stages.setdefault(st, {"kernels": {}})
        e["kernels"].setdefault(k, {})[c] = float(np.mean(vs))
    out = {"workload": line.get("config", line)["workload"], "copy_mib": copy_mib,
           "fetch_scale": scale.get("FETCH_SIZE"), "write_scale": scale.get("WRITE_SIZE"),
           "variants": {}, "stages": {}}
    for st, e in stages.items():
        # raw counters (upper case) add up over the stage's kernels; derived
        # ratios (LdsBankConflict, MeanOccupancyPerActiveCU) stay per kernel
        counters = sorted({c for kv in e["kernels"].values() for c in kv if c.isupper()})
        tot = {c: sum(kv.get(c, 0.0) for kv in e["kernels"].values()) for c in counters}
        rec = {"kernels": e["kernels"], "totals": tot}
        if "FETCH_SIZE" in tot and "WRITE_SIZE" in tot and len(scale) == 2:
            rec["fetch_bytes"] = tot["FETCH_SIZE"] * 1024.0 * scale["FETCH_SIZE"]
            rec["write_bytes"] = tot["WRITE_SIZE"] * 1024.0 * scale["WRITE_SIZE"]
            rec["bytes"] = rec["fetch_bytes"] + rec["write_bytes"]
            out["variants"][variants.get(st, st)] = rec["bytes"]
        out["stages"][variants.get(st, st)] = rec
    return out


def calibration_copy(dev, mib):
    """Three plain elementwise copies of ``mib`` MiB (read once, written
    once): the known-size kernel that scales the FETCH_SIZE / WRITE_SIZE
    counters of a profiled run to bytes."""
    n = mib * (1 << 20) // 4
    src = torch.ones(n, device=dev)
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize(dev)


def time_kernel(fn, reps, device):
    """Average duration of fn() from HIP events recorded on the launch stream.
    Python's cyclic garbage collector is run before and held off during the
    timed launches: freeing device objects of earlier work inside the timed
    region would stall the host between an event and its launch."""
    stream = torch.cuda.current_stream(device)
    fn()
    torch.cuda.synchronize(device)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    gc.collect()
    was_enabled = gc.isenabled()
    gc.disable()
    try:
        for s, e in ev:
            s.record(stream)
            fn()
            e.record(stream)
        torch.cuda.synchronize(device)
    finally:
        if was_enabled:
            gc.enable()
    return float(np.mean([s.elapsed_time(e) for s, e in ev]))


def cpu_baseline(wl, batch, frame=0, reps=5, warm=1, budget_s=30.0):
    """Oracle (C/OpenMP restatement of deepmr's CPU loops) on one frame of the
    workload: forward + adjoint of all coils; returns samples/s and the
    relative L2 of the GPU result for that frame against it."""
    from oracle import nufft_oracle as orc

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    lib = orc.load_loops(fast=True)
    lib.orc_set_threads(threads)
    torch.set_num_threads(threads)
    orc.use_fast_loops(True)
    orc.set_fft_backend("torch")
    try:
        n, ncoils = wl["n"], wl["ncoils"]
        coord = wl["coords"][frame]
        t0 = time.perf_counter()
        plan = orc.plan_nufft(coord, n, wl["width"], wl["osf"])
        t_plan = time.perf_counter() - t0
        x = (wl["smaps"] * wl["frames"][frame]).cpu().numpy()
        times = []
        y = xa = None
        for r in range(warm + reps):
            t0 = time.perf_counter()
            y = orc.apply_nufft(x, plan)
            xa = orc.apply_nufft_adj(y, plan)
            dt = time.perf_counter() - t0
            if r >= warm:
                times.append(dt)
            if sum(times) > budget_s:
                break
        med = float(np.median(times))
        nsamples = ncoils * coord.shape[0] * coord.shape[1]
        gpu_y = batch.kspace_view(frame).reshape(ncoils, -1).cpu().numpy()
        err = float(np.linalg.norm(gpu_y - y.reshape(ncoils, -1)) / np.linalg.norm(y))
        return dict(value=nsamples / med, unit="samples/s", cores=threads, kind="port",
                    sample=f"frame {frame} of the workload ({ncoils} coils x {coord.shape[0]}x"
                           f"{coord.shape[1]} samples) forward+adjoint, median of {len(times)} after "
                           f"{warm} warm-up, plan {t_plan:.2f} s excluded",
                    ms_per_step=med * 1e3), err
    finally:
        orc.use_fast_loops(False)
        orc.set_fft_backend("numpy")


def plain_radial_bench(dev, reps=10, warm=2):
    """Plain 2D single-coil NUFFT forward + adjoint: 128x128 image (complex
    N(0,1)), golden-angle radial 128 spokes x 256 samples in [-64, 64), osf
    2.0, KB width 4. Times the oracle's -ffast-math C/OpenMP restatement of
    deepmr's CPU loops with torch's CPU FFT on every host thread and on one
    thread (deepmr's gridding of one coil and frame runs on one thread:
    its parallel loop is over frame x coil), and the GPU path (plan reused)
    with HIP events. Medians after warm-up; plan time separate."""
    import tilenufft
    from oracle import nufft_oracle as orc
    from tilenufft import testing

    n = 128
    coord = testing.golden_angle_radial(128, 256, n)
    gen = torch.Generator().manual_seed(0)
    x = testing.complex_normal((1, n, n), generator=gen).numpy()
    nsamp = 128 * 256
    out = {"workload": "2D single-coil radial 128 spokes x 256, 128x128, osf 2.0, KB width 4, fwd+adj",
           "samples": nsamp}
    nproc = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    lib = orc.load_loops(fast=True)
    orc.use_fast_loops(True)
    orc.set_fft_backend("torch")
    try:
        t0 = time.perf_counter()
        plan = orc.plan_nufft(coord, n, 4, 2.0)
        out["cpu_plan_s"] = time.perf_counter() - t0
        for threads in (nproc, 1):
            lib.orc_set_threads(threads)
            torch.set_num_threads(threads)
            ts = []
            for r in range(warm + reps):
                t0 = time.perf_counter()
                y = orc.apply_nufft(x, plan)
                orc.apply_nufft_adj(y, plan)
                if r >= warm:
                    ts.append(time.perf_counter() - t0)
            med = float(np.median(ts))
            key = "cpu_1thread" if threads == 1 else "cpu"
            out[key] = {"ms": med * 1e3, "samples_per_s": nsamp / med, "cores": threads}
    finally:
        lib.orc_set_threads(nproc)
        torch.set_num_threads(nproc)
        orc.use_fast_loops(False)
        orc.set_fft_backend("numpy")
    gplan = tilenufft.fft.plan_nufft(coord, n, 4, 2.0, device=dev)
    xt = torch.from_numpy(x).to(dev)
    yk = tilenufft.fft.apply_nufft(xt, gplan)
    t = time_kernel(lambda: tilenufft.fft.apply_nufft_adj(tilenufft.fft.apply_nufft(xt, gplan), gplan), reps, dev)
    out["gpu"] = {"ms": t, "samples_per_s": nsamp / (t * 1e-3)}
    y_ref = orc.apply_nufft(x, orc.plan_nufft(coord, n, 4, 2.0))
    out["gpu_vs_oracle_rel_l2"] = float(np.linalg.norm(yk.cpu().numpy() - y_ref) / np.linalg.norm(y_ref))
    return out


def serving_bench(dev, nreq=48, rate_hz=400.0, seed=7, warm=6):
    """Mixed-size CG-SENSE stream (Poisson arrivals; N in {128,192,256,320},
    4..32 coils, radial or spiral, 10k..200k samples per coil, 10 CG
    iterations, osf 1.25, W 4) through the continuous-batching server and the
    same requests one at a time through the plain GPU NUFFT (plan reused over
    a request's iterations). Both runs see the same arrival times."""
    from tilenufft import serve

    def stream(n, s):
        return serve.make_stream(n, rate_hz, seed=s, device="cpu")

    # warm-up (kernel loading, rocFFT plans of the common sizes), untimed
    warm_srv = serve.ReconServer(dev, hbm_budget=64 << 30)
    warm_srv.run(stream(warm, seed + 1000), realtime=False)
    warm_srv.close()
    serve.run_plain(stream(warm, seed + 1000), dev, realtime=False)
    torch.cuda.synchronize(dev)

    reqs_s = stream(nreq, seed)
    srv = serve.ReconServer(dev, hbm_budget=64 << 30)
    srv.run(reqs_s, realtime=True)
    srv.close()
    reqs_p = stream(nreq, seed)
    serve.run_plain(reqs_p, dev, realtime=True)
    reqs_r = stream(nreq, seed)
    serve.run_plain(reqs_r, dev, realtime=True, replan=True)
    st_s, st_p, st_r = (serve.latency_stats(reqs_s), serve.latency_stats(reqs_p),
                        serve.latency_stats(reqs_r))
    errs = []
    for a, b in zip(reqs_s, reqs_p):
        if a.result is not None and b.result is not None:
            ra, rb = a.result.to(dev), b.result.to(dev)
            errs.append(float(torch.linalg.vector_norm(ra - rb) / torch.linalg.vector_norm(rb)))
    return dict(
        stream=f"{nreq} requests, Poisson {rate_hz:g} Hz (seed {seed}), N in {{128,192,256,320}}, "
               "4-32 coils, radial/spiral, 10k-200k samples/coil, 10 CG-SENSE iterations, osf 1.25, W 4",
        served=dict(p50_s=st_s["p50"], p99_s=st_s["p99"], recon_per_s=st_s["recon_per_s"],
                    completed=st_s["completed"], server_steps=srv.steps, max_active=srv.max_active,
                    admission=srv.admission, stats=srv.stats),
        plain=dict(p50_s=st_p["p50"], p99_s=st_p["p99"], recon_per_s=st_p["recon_per_s"],
                   completed=st_p["completed"],
                   path="one request at a time, apply_nufft / apply_nufft_adj with the request's plan "
                        "made once and reused over its CG iterations"),
        plain_per_call=dict(p50_s=st_r["p50"], p99_s=st_r["p99"], recon_per_s=st_r["recon_per_s"],
                            completed=st_r["completed"],
                            path="one request at a time through the functional nufft / nufft_adj, "
                                 "which plan on every call as deepmr's do"),
        speedup_recon_per_s=(st_s["recon_per_s"] / st_p["recon_per_s"]) if st_p["recon_per_s"] else None,
        speedup_vs_per_call=(st_s["recon_per_s"] / st_r["recon_per_s"]) if st_r["recon_per_s"] else None,
        max_rel_l2_served_vs_plain=max(errs) if errs else None,
    )


def kooshball_bench(dev, osfs=(2.0, 1.5), compare_runs=False, compare_tiles=True, reps=3, grid_mode=0,
                    tune=False):
    """3D kooshball kernels: 128^3 image, 8 coils, 20000 golden-means spokes x
    256 samples, KB width 4, osf 2.0 (256^3 grid) and 1.5 (192^3): the
    automatic plan's gridding (the gridding matrix) and staged interpolation,
    the whole forward and adjoint transforms, with
    ``compare_tiles`` the register-tile gridding and with ``compare_runs`` the
    register-run gridding too; with ``tune`` the autotuner's choice among the
    matrix and register-tile griddings (with their split thresholds) and the
    staged interpolation's 3D tiles, timed the same way. Durations from HIP
    events on the launch stream; algorithmic bytes as for the 2D workload.
    Returns a dict."""
    import tilenufft
    from tilenufft import testing
    from tilenufft.batch import BatchedNUFFT
    from tilenufft.fft.nufft import retune

    coord = testing.kooshball(20000, 256, 128)
    out = {"workload": "3D kooshball 128^3, 8 coils, 20000 spokes x 256 samples, KB width 4", "cases": []}
    gen = torch.Generator().manual_seed(0)
    for osf in osfs:
        log(f"bench: kooshball osf {osf}")
        t0 = time.perf_counter()
        plan = tilenufft.fft.plan_nufft(coord, 128, 4, osf, device=dev, grid_mode=grid_mode)
        torch.cuda.synchronize(dev)
        case = {"osf": osf, "grid": list(plan.os_shape), "plan_s": time.perf_counter() - t0,
                "plan_bytes": plan.interpolator.device_bytes()}
        b = BatchedNUFFT([plan], 8)
        b.kspace.copy_(testing.complex_normal(tuple(b.kspace.shape), generator=gen).to(dev))
        b_interp, b_grid = b.algorithmic_bytes()
        case["grid_alg_bytes"], case["interp_alg_bytes"] = b_grid, b_interp
        t = time_kernel(lambda: b.gridding(b.grid), reps, dev)
        case["grid_ms"] = t
        case["grid_frac"] = b_grid / (t * 1e-3) / 1e9 / HBM_PEAK_GBS
        ti = time_kernel(lambda: b.interp(b.grid), reps, dev)
        case["interp_ms"] = ti
        case["interp_frac"] = b_interp / (ti * 1e-3) / 1e9 / HBM_PEAK_GBS
        # whole transforms: forward (FFT stages + interpolation) and adjoint
        # (gridding + FFT stages) of the 8 coil images
        case["fft_stages"] = bool(plan.interpolator.has_fft)
        case["nufft_fwd_ms"] = time_kernel(lambda: b.forward(), reps, dev)
        case["nufft_adj_ms"] = time_kernel(lambda: b.adjoint(), reps, dev)
        b = None
        if compare_tiles:
            retune(plan, grid_mode=4)  # register tiles
            torch.cuda.synchronize(dev)
            case["tiles_plan_bytes"] = plan.interpolator.device_bytes()
            b = BatchedNUFFT([plan], 8)
            b.kspace.copy_(testing.complex_normal(tuple(b.kspace.shape), generator=gen).to(dev))
            case["grid_tiles_ms"] = time_kernel(lambda: b.gridding(b.grid), reps, dev)
            b = None
        if compare_runs:
            retune(plan, grid_mode=2, tile=(16, 16, 16))
            b = BatchedNUFFT([plan], 8)
            case["grid_runs_ms"] = time_kernel(lambda: b.gridding(b.grid), reps, dev)
            b = None
        if tune:
            from tilenufft.autotune import default_candidates, tune_batch

            # the fast 3D candidates only: the register-run and LDS-atomic
            # griddings take 25-300 ms here and are never chosen
            cand = dict(default_candidates(3), grid_modes=[3, 4], grid_chunks=[0], matrix_splits=[128, 256],
                        tiles_splits=[256, 384], interp_chunks=[], itiles=[(8, 8, 16), (8, 8, 8)],
                        staged_chunks=[2])
            log(f"bench: kooshball osf {osf} autotune")
            t0 = time.perf_counter()
            best = tune_batch([plan], 8, candidates=cand, reps=reps, use_cache=False)
            torch.cuda.synchronize(dev)
            tuned = {"tune_s": time.perf_counter() - t0, "grid_mode": best["grid_mode"],
                     "grid_split": best["grid_split"], "interp_mode": best["interp_mode"], "itile": best["itile"],
                     "interp_chunk": best["interp_chunk"], "plan_bytes": plan.interpolator.device_bytes()}
            b = BatchedNUFFT([plan], 8)
            b.kspace.copy_(testing.complex_normal(tuple(b.kspace.shape), generator=gen).to(dev))
            tuned["grid_ms"] = time_kernel(lambda: b.gridding(b.grid), reps, dev)
            tuned["grid_frac"] = b_grid / (tuned["grid_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS
            tuned["interp_ms"] = time_kernel(lambda: b.interp(b.grid), reps, dev)
            tuned["interp_frac"] = b_interp / (tuned["interp_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS
            tuned["nufft_fwd_ms"] = time_kernel(lambda: b.forward(), reps, dev)
            tuned["nufft_adj_ms"] = time_kernel(lambda: b.adjoint(), reps, dev)
            case["tuned"] = tuned
            b = None
        out["cases"].append(case)
        plan = None
        torch.cuda.empty_cache()
    return out


def cgsense_bench(steps=10, warmup=3):
    """Unrolled CG-SENSE with autograd: 256^2 image, 8 coils, 201 golden-angle
    radial spokes x 512 samples, KB width 4, osf 1.25 (320^2 grid), 10 CGStep
    iterations from zero, MSE loss against a seeded target, backward to the
    k-space. One step = forward + backward, timed eagerly and then with the
    module captured in HIP graphs (capture_unrolled); durations come from
    HIP events on the launch stream after three untimed steps. The encoding
    operator (plans) is built once, outside the timed steps."""
    import tilenufft
    from tilenufft import testing

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    n, nc, nspokes, nsamp, niter = 256, 8, 201, 512, 10
    coord = testing.golden_angle_radial(nspokes, nsamp, n)
    smaps = torch.from_numpy(testing.gaussian_coil_maps(nc, (n, n), seed=0)).to(dev)
    gen = torch.Generator().manual_seed(0)
    y = testing.complex_normal((nc, nspokes, nsamp), generator=gen).to(dev)
    target = (testing.complex_normal((n, n), generator=gen) * 0.1).to(dev)
    t0 = time.perf_counter()
    model = tilenufft.recon.UnrolledCGSENSE(torch.from_numpy(coord), (n, n), smaps, niter=niter, device=dev)
    torch.cuda.synchronize(dev)
    plan_s = time.perf_counter() - t0

    def step():
        yy = y.detach().requires_grad_(True)
        loss = (model(yy) - target).abs().pow(2).mean()
        loss.backward()

    def result():
        yy = y.detach().requires_grad_(True)
        x = model(yy)
        (x - target).abs().pow(2).mean().backward()
        return x.detach().clone(), yy.grad.detach().clone()

    log("bench: CG-SENSE 10 iterations, forward + backward (eager)")
    for _ in range(warmup):  # first steps grow the allocator's pools and lay out the autograd buffers
        step()
    torch.cuda.synchronize(dev)
    t_eager = time_kernel(step, steps, dev)
    x_ref, g_ref = result()
    log("bench: CG-SENSE captured in HIP graphs (forward and backward)")
    tilenufft.recon.capture_unrolled(model, y.detach().clone().requires_grad_(True))
    for _ in range(warmup):
        step()
    torch.cuda.synchronize(dev)
    t_graph = time_kernel(step, steps, dev)
    x_g, g_g = result()
    rel = lambda a, b: float((a - b).abs().norm() / b.abs().norm())
    out = {"workload": "unrolled CG-SENSE, 256^2, 8 coils, 201 radial spokes x 512, osf 1.25, W 4, "
                       "10 iterations, forward + backward (autograd) per step",
           "ms_per_step_eager": t_eager, "ms_per_step_graphed": t_graph,
           "ms_per_iteration_fwd_bwd_graphed": t_graph / niter, "plan_s": plan_s, "steps": steps,
           "warmup": warmup, "graphed_vs_eager_rel_l2": {"image": rel(x_g, x_ref), "grad": rel(g_g, g_ref)}}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-graph", action="store_true", help="replay eagerly instead of a HIP graph")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--kernel-reps", type=int, default=20)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-stream leg")
    ap.add_argument("--no-plain-radial", action="store_true", help="skip the plain single-coil 2D radial leg")
    ap.add_argument("--no-kooshball", action="store_true", help="skip the 3D kooshball kernel leg")
    ap.add_argument("--no-cgsense", action="store_true", help="skip the unrolled CG-SENSE (autograd) leg")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only replay the step and time the kernels (short runs under a profiler)")
    ap.add_argument("--serve-requests", type=int, default=48)
    ap.add_argument("--serve-rate", type=float, default=400.0)
    ap.add_argument("--tune-cache", default=None, help="directory of the autotuner's cache file")
    ap.add_argument("--grid-mode", type=int, default=0,
                    help="with --kooshball: gridding variant of the timed plan (0 automatic, 4 register tiles)")
    ap.add_argument("--grid-split", type=int, default=None,
                    help="after tuning, force this gridding split (for the matrix: the long-row threshold), "
                         "to profile each variant the tuner may pick")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="after measuring, run 3 plain device copies of this many MiB (a known-size "
                         "kernel that scales rocprofv3 FETCH_SIZE / WRITE_SIZE counters to bytes)")
    ap.add_argument("--fold-counters", nargs="+", default=None, metavar="DIR",
                    help="instead (no GPU): fold rocprofv3 --pmc output directories of --kernels-only runs into "
                         "per-stage bytes per launch; needs --bench-line and --calibrate-copy; writes --traffic-out")
    ap.add_argument("--bench-line", default=None, help="JSON line printed by one of the folded runs")
    ap.add_argument("--traffic-out", default=None, help="where --fold-counters writes its JSON (default: stdout)")
    ap.add_argument("--kooshball", action="store_true",
                    help="instead: time 3D kooshball gridding / interpolation (128^3, 8 coils, 20000 x 256, "
                         "W 4, osf 2.0 and 1.5) with register runs and the gridding matrix; one JSON line")
    ap.add_argument("--cgsense", action="store_true",
                    help="instead: time 10 unrolled CG-SENSE iterations forward + backward (256^2, 8 coils, "
                         "201 radial spokes x 512, osf 1.25); one JSON line")
    args = ap.parse_args()
    if args.fold_counters:
        if not args.bench_line or args.calibrate_copy <= 0:
            raise SystemExit("--fold-counters needs --bench-line and --calibrate-copy")
        res = fold_counters(args.fold_counters, args.bench_line, args.calibrate_copy)
        if args.traffic_out and os.path.exists(args.traffic_out):  # add to the variants already profiled
            with open(args.traffic_out) as f:
                old = json.load(f)
            for key in ("variants", "stages"):
                res[key] = {**old.get(key, {}), **res[key]}
        text = json.dumps(res, indent=1, sort_keys=True)
        if args.traffic_out:
            with open(args.traffic_out, "w") as f:
                f.write(text + "\n")
        print(text)
        return
    if args.kooshball:
        dev = torch.device("cuda:0")
        torch.cuda.set_device(dev)
        # --kernels-only: the automatic plan only
        res = kooshball_bench(dev, compare_runs=not args.kernels_only, compare_tiles=not args.kernels_only,
                              grid_mode=args.grid_mode, tune=not args.kernels_only)
        if args.calibrate_copy > 0:
            calibration_copy(dev, args.calibrate_copy)
        print(json.dumps(res), flush=True)
        return
    if args.cgsense:
        print(json.dumps(cgsense_bench()), flush=True)
        return
    if args.tune_cache:
        os.environ["TILENUFFT_CACHE"] = args.tune_cache
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the NUFFT path "
              "shards by frame, coil group or request with one process and one set of plans per GPU, and needs "
              "no communication except a per-iteration image all-reduce for coil-split CG-SENSE", flush=True)
        sys.exit(2)

    import tilenufft
    from tilenufft.batch import BatchedNUFFT

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    log("bench: building the spiral workload (8 frames x 16 coils, 256^2, osf 2)")
    wl = make_spiral_workload(dev)
    plans, ncoils = wl["plans"], wl["ncoils"]
    tune = None
    if not args.no_autotune:
        log("bench: autotuning tile / chunk sizes")
        tune = tilenufft.autotune.tune_batch(plans, ncoils)
        tune = {k: v for k, v in tune.items() if k != "trials"}
        log(f"bench: autotune -> {tune}")
    if args.grid_split is not None:
        from tilenufft.fft.nufft import retune

        for p in plans:
            retune(p, grid_split=args.grid_split)
    batch = BatchedNUFFT(plans, ncoils)
    for f in range(len(plans)):
        batch.image_view(f)[:, 0].copy_(wl["smaps"] * wl["frames"][f])

    step = (lambda: (batch.forward(), batch.adjoint()))
    graph = None
    if not args.no_graph:
        log("bench: capturing forward+adjoint into a HIP graph")
        graph = batch.capture()
        run = graph.replay
    else:
        run = step

    for _ in range(args.warmup):
        run()
    torch.cuda.synchronize(dev)
    log(f"bench: timing {args.steps} steps")
    gc.collect()
    gc.disable()  # no collector pauses between the timed replays
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run()
    torch.cuda.synchronize(dev)
    elapsed = time.perf_counter() - t0
    gc.enable()
    ms_step = elapsed / args.steps * 1e3

    # per-kernel durations on the launch stream (the forward output in the
    # k-space arena is kept for the parity check against the CPU baseline)
    kspace_saved = batch.kspace.clone()
    kgrid = torch.fft.fftn(batch.grid, dim=(-2, -1))
    t_interp = time_kernel(lambda: batch.interp(kgrid), args.kernel_reps, dev)
    batch.kspace.copy_(kspace_saved)
    gout = torch.empty_like(batch.grid)
    t_grid = time_kernel(lambda: batch.gridding(gout), args.kernel_reps, dev)
    t_copy = time_kernel(lambda: batch.image_adj.copy_(batch.image), args.kernel_reps, dev)
    # pruned FFT stages (image <-> grid): algorithmic bytes = image read or
    # written once + grid written or read once
    native0 = plans[0].interpolator
    t_fft = {}
    b_fft = 0
    if native0.has_fft:
        t_fft["fft_fwd"] = time_kernel(lambda: native0.fft_forward(batch.image, gout, batch.nimg),
                                       args.kernel_reps, dev)
        t_fft["fft_adj"] = time_kernel(lambda: native0.fft_adjoint(gout, batch.image_adj, batch.nimg),
                                       args.kernel_reps, dev)
        b_fft = 8 * batch.nimg * (int(np.prod(plans[0].shape)) + int(np.prod(plans[0].os_shape)))
    batch.kspace.copy_(kspace_saved)
    b_interp, b_grid = batch.algorithmic_bytes()
    gbs_interp = b_interp / (t_interp * 1e-3) / 1e9
    gbs_grid = b_grid / (t_grid * 1e-3) / 1e9
    times = dict(grid=t_grid, interp=t_interp, **t_fft)
    dominant = max(times, key=times.get)
    ach = {"grid": gbs_grid, "interp": gbs_interp}.get(dominant, b_fft / (times[dominant] * 1e-3) / 1e9)
    # HBM bytes per launch of the timed variant, from rocprofv3 FETCH_SIZE /
    # WRITE_SIZE runs of this workload (scaled by a copy kernel's counters);
    # null when the autotuner picked a variant that was not profiled
    variant = variant_key(plans[0], dominant, ncoils)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_spiral2d.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("variants", {}).get(variant)

    total_samples = batch.nsamples
    value = total_samples / (ms_step * 1e-3)

    cpu = None
    parity = None
    if not args.no_cpu_baseline and not args.kernels_only:
        log("bench: CPU baseline (oracle restatement of deepmr's CPU loops) on frame 0")
        cpu, parity = cpu_baseline(wl, batch)

    if args.calibrate_copy > 0:
        calibration_copy(dev, args.calibrate_copy)

    plain_radial = None
    if not args.no_plain_radial and not args.kernels_only:
        log("bench: plain single-coil 2D radial (CPU oracle on all threads / one thread, GPU)")
        plain_radial = plain_radial_bench(dev)
    kooshball = None
    if not args.no_kooshball and not args.kernels_only:
        kooshball = kooshball_bench(dev, tune=True)

    cgsense = None
    if not args.no_cgsense and not args.kernels_only:
        log("bench: unrolled CG-SENSE with autograd (256^2, 8 coils), eager and graphed")
        cgsense = cgsense_bench()

    serving = None
    if not args.no_serve and not args.kernels_only:
        log(f"bench: serving stream ({args.serve_requests} requests) served vs plain")
        serving = serving_bench(dev, args.serve_requests, args.serve_rate)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {
            "workload": "batched 2D multicoil VD spiral, 8 frames x 16 coils, 256x256, osf 2.0, "
                        "KB width 4, 24-48 interleaves x 2048 samples per frame, fwd+adj of all "
                        "frames/coils per step" + (" in one HIP graph" if graph is not None else ""),
            "frames": len(plans),
            "coils": ncoils,
            "interleaves": wl["interleaves"],
            "samples_per_step": total_samples,
            "autotune": tune,
            "hip_graph": graph is not None,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dominant,
            "achieved": ach,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": ach / HBM_PEAK_GBS,
            "traffic": traffic,
            "variant": variant,
        },
        "kernels": {
            "interp_ms": t_interp,
            "grid_ms": t_grid,
            "interp_alg_bytes": b_interp,
            "grid_alg_bytes": b_grid,
            "interp_GBs": gbs_interp,
            "grid_GBs": gbs_grid,
            "interp_frac": gbs_interp / HBM_PEAK_GBS,
            "grid_frac": gbs_grid / HBM_PEAK_GBS,
            "image_copy_ms": t_copy,
            **{f"{k}_ms": v for k, v in t_fft.items()},
            **({"fft_alg_bytes": b_fft} if t_fft else {}),
            "variants": {k: variant_key(plans[0], k, ncoils) for k in ("grid", "interp", *t_fft)},
            **{f"{k}_frac": b_fft / (v * 1e-3) / 1e9 / HBM_PEAK_GBS for k, v in t_fft.items()},
        },
        "cpu_baseline": cpu,
        "parity_rel_l2_frame0_vs_oracle": parity,
        "plain_radial_2d": plain_radial,
        "kooshball_3d": kooshball,
        "cgsense_2d": cgsense,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
