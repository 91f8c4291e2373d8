"""Autotuner for the launch parameters of the gridding / interpolation kernels.

Tunable per plan: the gridding variant and tile (bin size), the coils per
gridding workgroup, the hot-tile split threshold, and the interpolation
variant, tile and coils per work item. Candidates are timed on
the actual batch with HIP events; the choice is cached on disk keyed by the
geometry (ndim, N, G, W, coils, samples per trajectory rounded, number of
trajectories).
"""

__all__ = ["tune_batch", "default_candidates"]

import json
import os

import numpy as np
import torch

from .batch import BatchedNUFFT
from .fft.nufft import retune


def default_candidates(ndim):
    if ndim == 2:
        tiles = [(32, 16), (16, 32), (32, 32), (16, 16), (16, 64), (8, 32)]
        itiles = [(16, 16), (32, 32), (16, 64), (32, 64)]
    else:
        tiles = [(8, 8, 8), (8, 16, 16), (16, 16, 16), (8, 8, 32), (16, 16, 32)]
        itiles = [(4, 8, 8), (8, 8, 8), (8, 8, 16), (8, 16, 16)]
    # register tiles (grid_mode 4) are 3D only; plans they do not fit skip them
    return dict(tiles=tiles, grid_chunks=[0, 1, 2, 4, 8, 16], interp_chunks=[1, 2, 4, 8, 16],
                itiles=itiles, staged_chunks=[2, 4, 8, 16], grid_splits=[512, 1024, 2048, 4096],
                matrix_splits=[64, 128, 256], tiles_splits=[256, 384, 512, 1024],
                grid_modes=[3, 2, 1] if ndim == 2 else [3, 4, 2, 1])


def _cache_path():
    root = os.environ.get("TILENUFFT_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "tilenufft"))
    return os.path.join(root, "autotune.json")


def _key(plans, ncoils):
    p = plans[0]
    m = int(np.median([q.npts * q.nframes for q in plans]))
    bucket = int(2 ** round(np.log2(max(m, 1))))
    # the library is built for gfx950 only, so the geometry alone identifies a
    # tuning problem (device-name queries are not stable under profilers)
    return f"gfx950|d{p.ndim}|N{p.shape}|G{p.os_shape}|W{p.width}|C{max(ncoils)}|M{bucket}|P{len(plans)}"


def _time(fn, reps, device):
    torch.cuda.synchronize(device)
    fn()
    start = [torch.cuda.Event(enable_timing=True) for _ in range(reps)]
    end = [torch.cuda.Event(enable_timing=True) for _ in range(reps)]
    for s, e in zip(start, end):
        s.record()
        fn()
        e.record()
    torch.cuda.synchronize(device)
    return float(np.median([s.elapsed_time(e) for s, e in zip(start, end)]))


def _apply(plans, tile, ic, gc, mode=0, itile=None, split=0, gmode=None):
    for p in plans:
        retune(p, tile=tile, interp_chunk=ic, grid_chunk=gc, interp_mode=mode, itile=itile,
               grid_split=split, grid_mode=gmode)


def tune_batch(plans, ncoils, candidates=None, reps=5, use_cache=True, verbose=False):
    """Pick tile / chunk sizes for ``plans`` (rebuilt in place) and return them.

    Returns dict(tile, grid_chunk, interp_mode, itile, interp_chunk, grid_ms, interp_ms).
    """
    ncoils = [int(ncoils)] * len(plans) if np.isscalar(ncoils) else list(ncoils)
    key = _key(plans, ncoils)
    path = _cache_path()
    cache = {}
    if use_cache and os.path.exists(path):
        try:
            with open(path) as f:
                cache = json.load(f)
        except (OSError, ValueError):
            cache = {}
    if use_cache and key in cache and "grid_mode" in cache[key]:
        best = cache[key]
        _apply(plans, tuple(best["tile"]), best["interp_chunk"], best["grid_chunk"],
               best.get("interp_mode", 0), tuple(best["itile"]) if best.get("itile") else None,
               best.get("grid_split", 0), best["grid_mode"])
        return dict(best, cached=True)

    dev = plans[0].interpolator.device
    cand = candidates or default_candidates(plans[0].ndim)
    results = []
    with torch.cuda.device(dev):
        # 1) gridding variant x tile, automatic chunks
        best_tile, best_t, best_gm = None, float("inf"), 0
        for gm in cand.get("grid_modes", [0]):
            # the matrix and register-tile griddings do not use the gridding tile
            for tile in (cand["tiles"][:1] if gm in (3, 4) else cand["tiles"]):
                try:
                    _apply(plans, tile, 0, 0, gmode=gm)
                    b = BatchedNUFFT(plans, ncoils)
                    t = _time(lambda: b.gridding(b.grid), reps, dev)
                except RuntimeError as exc:  # variant unsupported for this width
                    results.append(("tile", (gm, tile), str(exc)))
                    continue
                results.append(("tile", (gm, tile), t))
                if verbose:
                    print(f"autotune grid_mode {gm} tile {tile}: grid {t:.3f} ms", flush=True)
                if t < best_t:
                    best_tile, best_t, best_gm = tile, t, gm
        # 2) coils per gridding workgroup
        best_gc, best_gt = 0, best_t
        for gc in cand["grid_chunks"]:
            if gc == 0:
                continue
            _apply(plans, best_tile, 0, gc, gmode=best_gm)
            b = BatchedNUFFT(plans, ncoils)
            t = _time(lambda: b.gridding(b.grid), reps, dev)
            results.append(("grid_chunk", gc, t))
            if t < best_gt:
                best_gc, best_gt = gc, t
        # 2b) hot-tile split threshold
        best_split = 0
        # (for the matrix gridding: the row length above which a whole
        # workgroup sums one cell)
        splits = "matrix_splits" if best_gm == 3 else "tiles_splits" if best_gm == 4 else "grid_splits"
        for sp in cand.get(splits, []):
            _apply(plans, best_tile, 0, best_gc, split=sp, gmode=best_gm)
            b = BatchedNUFFT(plans, ncoils)
            t = _time(lambda: b.gridding(b.grid), reps, dev)
            results.append(("grid_split", sp, t))
            if verbose:
                print(f"autotune grid_split {sp}: grid {t:.3f} ms", flush=True)
            if t < best_gt:
                best_split, best_gt = sp, t
        # 3) interpolation: gather variant (coils per item) and LDS-staged
        #    variant (tile x coils per item)
        interp_cands = [(1, None, ic) for ic in [0] + list(cand["interp_chunks"])]
        interp_cands += [(2, it, ic) for it in cand.get("itiles", []) for ic in cand.get("staged_chunks", [])]
        best_i, best_it = None, float("inf")
        for mode, itile, ic in interp_cands:
            try:
                _apply(plans, best_tile, ic, best_gc, mode, itile, best_split, best_gm)
                b = BatchedNUFFT(plans, ncoils)
                t = _time(lambda: b.interp(b.grid), reps, dev)
            except RuntimeError as exc:  # e.g. a tile too large for LDS
                results.append(("interp", (mode, itile, ic), str(exc)))
                continue
            results.append(("interp", (mode, itile, ic), t))
            if verbose:
                print(f"autotune interp mode={mode} itile={itile} chunk={ic}: {t:.3f} ms", flush=True)
            if t < best_it:
                best_i, best_it = (mode, itile, ic), t
        mode, itile, best_ic = best_i
        _apply(plans, best_tile, best_ic, best_gc, mode, itile, best_split, best_gm)
    best = dict(tile=list(best_tile), grid_mode=best_gm, grid_chunk=best_gc, grid_split=best_split,
                interp_mode=mode,
                itile=list(itile) if itile else None, interp_chunk=best_ic, grid_ms=best_gt,
                interp_ms=best_it)
    if use_cache:
        cache[key] = best
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            with open(path, "w") as f:
                json.dump(cache, f, indent=1)
        except OSError:
            pass
    return dict(best, cached=False, trials=results)
