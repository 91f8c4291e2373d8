"""Batched NUFFT over many trajectories with uneven sample counts.

deepmr takes one trajectory set per call and launches one kernel per call
(src/deepmr/fft/_interp/interp.py:682-742, grid.py:1073-1166). Here every
(trajectory, coil) pair of a batch shares single launches:

  forward  one apod+pad launch -> one batched rocFFT -> one interpolation
           launch over a flattened (trajectory, sample block, coil range) queue
  adjoint  one gridding launch over a (trajectory, tile, coil range) queue ->
           one batched rocFFT -> one crop+apod launch

All trajectories of a batch share the image geometry (N, G, W, osf); the
sample count may differ per trajectory. Buffers are static arenas so the
forward+adjoint pair can be captured in a HIP graph and replayed.
"""

__all__ = ["BatchedNUFFT"]

import ctypes

import numpy as np
import torch

from . import _core, _lib


class BatchedNUFFT:
    """Forward / adjoint NUFFT of several plans in shared launches.

    plans   : list of NUFFTPlan (same shape / os_shape / width / oversamp)
    ncoils  : coils per plan (int or list)
    Images are laid out per plan as (ncoils, nframes, *N) (nframes = the
    plan's frames), k-space as (ncoils, nframes, npts); see ``image_view`` /
    ``kspace_view`` for per-plan views into the arenas.
    """

    def __init__(self, plans, ncoils, device=None):
        if not plans:
            raise ValueError("need at least one plan")
        p0 = plans[0]
        for p in plans[1:]:
            if (p.shape, p.os_shape, p.width, p.ndim) != (p0.shape, p0.os_shape, p0.width, p0.ndim):
                raise ValueError("all plans of a batch must share shape, os_shape and width")
            if tuple(np.float32(p.oversamp)) != tuple(np.float32(p0.oversamp)):
                raise ValueError("all plans of a batch must share oversamp")
        self.plans = list(plans)
        self.ncoils = [int(ncoils)] * len(plans) if np.isscalar(ncoils) else [int(c) for c in ncoils]
        if len(self.ncoils) != len(plans):
            raise ValueError("ncoils must have one entry per plan")
        self.device = p0.interpolator.device if device is None else _core.compute_device(device)
        self.ndim = p0.ndim
        self.shape = tuple(p0.shape)
        self.os_shape = tuple(p0.os_shape)
        gvol = int(np.prod(self.os_shape))
        nvol = int(np.prod(self.shape))
        self.img_off, self.grid_off, self.samp_off = [], [], []
        gi = si = ii = 0
        for p, c in zip(self.plans, self.ncoils):
            self.img_off.append(ii)
            self.grid_off.append(gi)
            self.samp_off.append(si)
            ii += c * p.nframes * nvol
            gi += c * p.nframes * gvol
            si += c * p.nframes * p.npts
        self.nimg = sum(c * p.nframes for p, c in zip(self.plans, self.ncoils))
        self.nsamples = si
        self.image = torch.zeros((self.nimg, *self.shape), dtype=torch.complex64, device=self.device)
        self.image_adj = torch.zeros_like(self.image)
        self.grid = torch.zeros((self.nimg, *self.os_shape), dtype=torch.complex64, device=self.device)
        self.kspace = torch.zeros((max(si, 1),), dtype=torch.complex64, device=self.device)
        self._handle = None
        self._build()

    # -- native batch tables ------------------------------------------------
    def _build(self):
        lib = _lib.load()
        n = len(self.plans)
        handles = (ctypes.c_void_p * n)(*[p.interpolator.handle.value for p in self.plans])
        nb = (ctypes.c_int64 * n)(*self.ncoils)
        go = (ctypes.c_int64 * n)(*self.grid_off)
        so = (ctypes.c_int64 * n)(*self.samp_off)
        out = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(lib.tn_batch_create(handles, nb, go, so, None, n, ctypes.byref(out)),
                       "tn_batch_create")
        self._handle = out
        import weakref

        self._fin = weakref.finalize(self, _destroy_batch, out.value, self.device)

    def image_view(self, k, adjoint=False):
        """Image arena slice of plan k, shaped (ncoils, nframes, *N): the
        forward input, or the adjoint output with ``adjoint=True``."""
        p, c = self.plans[k], self.ncoils[k]
        n = c * p.nframes
        start = self.img_off[k] // int(np.prod(self.shape))
        arena = self.image_adj if adjoint else self.image
        return arena[start:start + n].view(c, p.nframes, *self.shape)

    def kspace_view(self, k):
        """k-space arena slice of plan k, shaped (ncoils, nframes, npts)."""
        p, c = self.plans[k], self.ncoils[k]
        n = c * p.nframes * p.npts
        return self.kspace[self.samp_off[k]:self.samp_off[k] + n].view(c, p.nframes, p.npts)

    # -- transforms on the arenas --------------------------------------------
    def forward(self):
        """image arena -> k-space arena (deepmr apply_nufft per plan)."""
        native0 = self.plans[0].interpolator
        if native0.has_fft:
            k = native0.fft_forward(self.image, self.grid, self.nimg)
        else:
            native0.apod_pad(self.image, self.grid, self.nimg)
            k = torch.fft.fftn(self.grid, dim=tuple(range(-self.ndim, 0)))
        self.interp(k)
        return self.kspace

    def adjoint(self):
        """k-space arena -> adjoint image arena (deepmr apply_nufft_adj per plan)."""
        self.gridding(self.grid)
        native0 = self.plans[0].interpolator
        if native0.has_fft:
            native0.fft_adjoint(self.grid, self.image_adj, self.nimg)
        else:
            img = torch.fft.ifftn(self.grid, dim=tuple(range(-self.ndim, 0)))
            native0.crop_apod(img, self.image_adj, self.nimg)
        return self.image_adj

    def interp(self, grid):
        """Interpolation launch only: grid arena layout -> k-space arena."""
        _core.check_tensor(grid, self.device, "grid")
        _lib.check(_lib.load().tn_batch_interp(self._handle, _core.ptr(grid), _core.ptr(self.kspace),
                                               None, _core.stream_ptr(self.device)), "tn_batch_interp")

    def gridding(self, grid_out):
        """Gridding launch only: k-space arena -> grid arena layout."""
        _core.check_tensor(grid_out, self.device, "grid")
        _lib.check(_lib.load().tn_batch_grid(self._handle, _core.ptr(self.kspace), _core.ptr(grid_out),
                                             None, _core.stream_ptr(self.device)), "tn_batch_grid")

    # -- HIP graph ------------------------------------------------------------
    def capture(self, fn=None):
        """Capture ``fn`` (default: forward then adjoint) into a HIP graph.

        Returns the torch CUDAGraph (HIP graph on ROCm); call ``.replay()``.
        The arenas are static: a replay reads ``image`` as left by the caller
        and overwrites ``kspace`` and ``image_adj``.
        """
        fn = fn or (lambda: (self.forward(), self.adjoint()))
        with torch.cuda.device(self.device):
            s = torch.cuda.Stream(self.device)
            s.wait_stream(torch.cuda.current_stream(self.device))
            with torch.cuda.stream(s):
                fn()  # warm-up: FFT plans and any lazy state outside capture
            torch.cuda.current_stream(self.device).wait_stream(s)
            torch.cuda.synchronize(self.device)
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph):
                fn()
        return graph

    # -- accounting -------------------------------------------------------------
    def algorithmic_bytes(self):
        """(interp, grid) algorithmic HBM bytes per launch.

        interp: 8*C*M (samples written) + 8*C*G_fp (touched grid read once)
                + 4*d*M (coordinates read once)
        grid:   8*C*M (samples read) + 8*C*G^d (full grid written once)
                + 4*d*M
        summed over frames; G_fp is the union of the kernel footprints.
        """
        gvol = int(np.prod(self.os_shape))
        bi = bg = 0
        for p, c in zip(self.plans, self.ncoils):
            m = p.nframes * p.npts
            fp = p.interpolator.footprint_cells()
            bi += 8 * c * m + 8 * c * fp + 4 * self.ndim * m
            bg += 8 * c * m + 8 * c * gvol * p.nframes + 4 * self.ndim * m
        return bi, bg


def _destroy_batch(handle, device):
    try:
        with torch.cuda.device(device):
            torch.cuda.current_stream(device).synchronize()
            _lib.load().tn_batch_destroy(ctypes.c_void_p(handle))
    except Exception:
        pass
