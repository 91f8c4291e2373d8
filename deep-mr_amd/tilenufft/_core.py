"""Host-side plan math and the device plan handle.

The plan arithmetic (Kaiser-Bessel beta, Cartesian-axis detection, oversampled
shape, coordinate scaling, kernel scaling, apodization) is computed on the
host exactly as deepmr does it (float32 where deepmr is float32), and the
device plan is then built by the native library.
"""

import ctypes
import math
import weakref

import numpy as np
import torch

from . import _lib

# --------------------------------------------------------------------------
# devices and streams
# --------------------------------------------------------------------------


def compute_device(device=None):
    """The HIP device that computes for a deepmr-style ``device`` argument.

    deepmr's ``device="cpu"`` (its default) selects the CPU Numba path; this
    package computes on HIP devices only, so "cpu" (and ``None``) map to the
    current HIP device while results still return on the caller's device.
    """
    if not torch.cuda.is_available():
        raise RuntimeError(
            "tilenufft computes on a HIP device (MI355X); torch.cuda.is_available() is False"
        )
    if device is None:
        return torch.device("cuda", torch.cuda.current_device())
    d = torch.device(device)
    if d.type == "cpu":
        return torch.device("cuda", torch.cuda.current_device())
    if d.type != "cuda":
        raise ValueError(f"unsupported device {device!r}")
    if d.index is None:
        d = torch.device("cuda", torch.cuda.current_device())
    return d


def stream_ptr(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def ptr(t):
    """Raw device pointer of ``t`` for the C ABI. Lazy conjugate / negative
    views are refused: their memory does not hold the values they denote."""
    if t is None:
        return ctypes.c_void_p(0)
    if t.is_conj() or t.is_neg():
        raise ValueError("resolve conjugate/negative views (resolve_conj/resolve_neg) before a native call")
    return ctypes.c_void_p(t.data_ptr())


# --------------------------------------------------------------------------
# plan arithmetic (src/deepmr/fft/nufft.py:84-119, :407-447;
#                  src/deepmr/fft/_interp/plan.py:190-248)
# --------------------------------------------------------------------------


def kaiser_bessel_beta(width, oversamp):
    """Beatty et al. (2005) beta = pi*sqrt((W/osf*(osf-1/2))^2 - 0.8), float32."""
    w = np.asarray(width, dtype=np.int16)
    o = np.asarray(oversamp, dtype=np.float32)
    inner = ((w / o) * (o - np.float32(0.5))) ** 2 - np.float32(0.8)
    return (np.float32(math.pi) * np.sqrt(inner.astype(np.float32))).astype(np.float32)


def kb_value(u, beta):
    """I0(beta*sqrt(1-u^2)) via Abramowitz & Stegun 9.8.1 / 9.8.2 (float64)."""
    if abs(u) > 1:
        return 0.0
    x = float(beta) * math.sqrt(1.0 - u * u)
    t = x / 3.75
    if x < 3.75:
        t2 = t * t
        return 1 + t2 * (3.5156229 + t2 * (3.0899424 + t2 * (1.2067492 + t2 * (
            0.2659732 + t2 * (0.0360768 + t2 * 0.0045813)))))
    q = 1.0 / t
    poly = 0.39894228 + q * (0.01328592 + q * (0.00225319 + q * (-0.00157565 + q * (
        0.00916281 + q * (-0.02057706 + q * (0.02635537 + q * (-0.01647633 + q * 0.00392377)))))))
    return poly * math.exp(x) / math.sqrt(x)


def kernel_scale(width, beta):
    """Sum of the taps of a sample sitting on a grid node, multiplied over axes.

    Width-1 (Cartesian) axes contribute 1. Axes in any consistent order.
    """
    s = 1.0
    for w, b in zip(width, beta):
        w = int(w)
        if w == 1:
            continue
        first = math.ceil(-w / 2)
        s *= sum(kb_value((first + k) / (w / 2), b) for k in range(w))
    return s


def cartesian_axes(coord, shape_zyx):
    """Axes sampled on integer positions, as deepmr decides it.

    deepmr indexes ``shape`` in (z, y, x) order but ``coord[..., ax]`` in
    (x, y, z) order (src/deepmr/fft/nufft.py:105-109); that pairing is kept so
    that anisotropic plans classify axes identically. Result in (z, y, x).
    """
    ndim = coord.shape[-1]
    flags = []
    for ax in range(ndim):
        a = (np.float32(shape_zyx[ax]) * coord[..., ax]).astype(np.float32)
        flags.append(bool(np.allclose(a, np.round(a))))
    return np.asarray(flags[::-1], dtype=bool)


def oversampled_shape(shape, oversamp):
    return np.ceil(np.asarray(oversamp, np.float32) * np.asarray(shape)).astype(np.int64)


def scale_coordinates(coord, shape_zyx, oversamp_zyx):
    """coord * G/N + G//2 per axis in float32 (coord last axis is x, y, z)."""
    out = np.array(coord, dtype=np.float32, copy=True)
    ndim = out.shape[-1]
    for d in range(ndim):  # d = 0 is x
        ax = ndim - 1 - d
        g = np.ceil(np.float32(oversamp_zyx[ax]) * np.float32(shape_zyx[ax]))
        factor = np.float32(g / np.float32(shape_zyx[ax]))
        out[..., d] = (out[..., d] * factor).astype(np.float32)
        out[..., d] = (out[..., d] + np.float32(g // 2)).astype(np.float32)
    return out


def apodization(n, oversamp, width, beta):
    """Deapodization weights along one axis (float32), normalised at n//2."""
    g = np.float32(np.ceil(np.float32(oversamp) * n))
    i = np.arange(n, dtype=np.float32) - np.float32(n // 2)
    arg = np.float32(beta) ** 2 - (np.float32(math.pi * int(width)) * i / g) ** 2
    with np.errstate(invalid="ignore", divide="ignore"):
        s = np.sqrt(arg.astype(np.float32))
        a = (s / np.sinh(s)).astype(np.float32)
        a = (a / a[n // 2]).astype(np.float32)
    return np.nan_to_num(a, nan=1.0).astype(np.float32)


# --------------------------------------------------------------------------
# device plan
# --------------------------------------------------------------------------


class NativePlan:
    """Owner of one ``tn_plan`` (device-resident sorted trajectory + tables).

    Arguments are in (z, y, x) order; ``coord`` is (nframes, npts, ndim) in
    grid units (x first). ``shift`` selects the output index convention
    (G - G//2 for the fused NUFFT, 0 for the plain interpolator).
    """

    def __init__(self, coord, shape, grid, width, beta, inv_scale, shift, device,
                 apod=None, tile=None, interp_chunk=0, grid_chunk=0, itile=None, interp_mode=0,
                 grid_split=0, grid_mode=0):
        lib = _lib.load()
        coord = np.ascontiguousarray(coord, dtype=np.float32)
        if coord.ndim != 3:
            raise ValueError("coord must be (nframes, npts, ndim)")
        nframes, npts, ndim = coord.shape
        if not np.isfinite(coord).all():
            raise ValueError("coordinates contain NaN or Inf")
        if npts and np.abs(coord).max() > 1e7:
            raise ValueError("coordinates out of range")
        self.ndim, self.nframes, self.npts = ndim, nframes, npts
        self.device = device
        self.shape = tuple(int(s) for s in shape)
        self.grid = tuple(int(g) for g in grid)
        self.width = tuple(int(w) for w in width)
        if ndim == 1:
            # a 1D transform is the 2D one of a unit y axis: one grid row,
            # width 1 (the Cartesian path: one tap of weight 1 at y = 0), the
            # same x taps, scale and apodization (deepmr's kernel scaling and
            # apodization give a width-1 axis the factor 1)
            coord = np.ascontiguousarray(np.concatenate([coord, np.zeros_like(coord)], -1))
            shape, grid, width = (1, *self.shape), (1, *self.grid), (1, *self.width)
            beta, shift = (0.0, *beta), (0, *shift)
            tile = None if tile is None else (1, *tile)
            itile = None if itile is None else (1, *itile)
            apod = None if apod is None else (None, *apod)
            ndim = 2
        pad = 3 - ndim
        p = _lib.PlanParams()
        p.ndim, p.nframes, p.npts = ndim, nframes, npts
        p.shape[:] = [1] * pad + [int(s) for s in shape]
        p.grid[:] = [1] * pad + [int(g) for g in grid]
        p.width[:] = [1] * pad + [int(w) for w in width]
        p.beta[:] = [0.0] * pad + [float(b) for b in beta]
        p.shift[:] = [0] * pad + [int(s) for s in shift]
        p.tile[:] = [0, 0, 0] if tile is None else [1] * pad + [int(t) for t in tile]
        p.itile[:] = [0, 0, 0] if itile is None else [1] * pad + [int(t) for t in itile]
        p.interp_mode = int(interp_mode)
        p.interp_chunk = int(interp_chunk)
        p.grid_chunk = int(grid_chunk)
        p.grid_split = int(grid_split)
        p.grid_mode = int(grid_mode)
        p.inv_scale = float(inv_scale)
        keep = []
        for d in range(ndim):
            if apod is not None and apod[d] is not None:
                a = np.ascontiguousarray(apod[d], dtype=np.float32)
                keep.append(a)
                p.apod[pad + d] = a.ctypes.data
        self.params = p
        with torch.cuda.device(device):
            dcoord = torch.from_numpy(coord.reshape(-1)).to(device)
            handle = ctypes.c_void_p()
            _lib.check(lib.tn_plan_create(ctypes.byref(p), ptr(dcoord), stream_ptr(device),
                                          ctypes.byref(handle)), "tn_plan_create")
            torch.cuda.current_stream(device).synchronize()
        del keep
        self.handle = handle
        self._finalizer = weakref.finalize(self, _destroy, handle.value, device)

    # -- introspection -----------------------------------------------------
    def device_bytes(self):
        return int(_lib.load().tn_plan_device_bytes(self.handle))

    def entries(self):
        return int(_lib.load().tn_plan_entries(self.handle))

    def footprint_cells(self):
        with torch.cuda.device(self.device):
            n = int(_lib.load().tn_plan_footprint_cells(self.handle))
        if n < 0:
            _lib.check(n, "tn_plan_footprint_cells")
        return n

    def grid_mode_effective(self):
        """The gridding variant the library chose (1, 2 or 3)."""
        m = int(_lib.load().tn_plan_grid_mode(self.handle))
        if m < 0:
            _lib.check(m, "tn_plan_grid_mode")
        return m

    def tile(self):
        t = (ctypes.c_int32 * 3)()
        _lib.check(_lib.load().tn_plan_tile(self.handle, t), "tn_plan_tile")
        return tuple(t)[3 - self.ndim:]

    def prepare(self, nbatch):
        with torch.cuda.device(self.device):
            _lib.check(_lib.load().tn_plan_prepare(self.handle, int(nbatch)), "tn_plan_prepare")

    # -- kernels (tensors must be contiguous complex64 on self.device) ------
    def interp(self, grid, out, nbatch, weight=None):
        _lib.check(_lib.load().tn_interp(self.handle, ptr(grid), ptr(out), int(nbatch), ptr(weight),
                                         stream_ptr(self.device)), "tn_interp")
        return out

    def gridding(self, samp, out, nbatch, weight=None):
        _lib.check(_lib.load().tn_grid(self.handle, ptr(samp), ptr(out), int(nbatch), ptr(weight),
                                       stream_ptr(self.device)), "tn_grid")
        return out

    def interp_lowrank(self, grid, out, nbatch, ncoeff, frame_npts, basis_adjoint, weight=None):
        """grid (nbatch, ncoeff, *G) -> out (nbatch, npts); one-frame plan."""
        _lib.check(_lib.load().tn_interp_lowrank(self.handle, ptr(grid), ptr(out), int(nbatch), int(ncoeff),
                                                 int(frame_npts), ptr(basis_adjoint), ptr(weight),
                                                 stream_ptr(self.device)), "tn_interp_lowrank")
        return out

    def gridding_lowrank(self, samp, out, nbatch, ncoeff, frame_npts, basis, weight=None):
        """samp (nbatch, npts) -> out (nbatch, ncoeff, *G); one-frame plan."""
        _lib.check(_lib.load().tn_grid_lowrank(self.handle, ptr(samp), ptr(out), int(nbatch), int(ncoeff),
                                               int(frame_npts), ptr(basis), ptr(weight),
                                               stream_ptr(self.device)), "tn_grid_lowrank")
        return out

    def apod_pad(self, image, out, nimg):
        _lib.check(_lib.load().tn_apod_pad(self.handle, ptr(image), ptr(out), int(nimg),
                                           stream_ptr(self.device)), "tn_apod_pad")
        return out

    def crop_apod(self, grid, out, nimg):
        _lib.check(_lib.load().tn_crop_apod(self.handle, ptr(grid), ptr(out), int(nimg),
                                            stream_ptr(self.device)), "tn_crop_apod")
        return out

    @property
    def has_fft(self):
        """True when the plan has the pruned, fused FFT stages (2D and 3D,
        every grid length among 16, 24, 32, 40, 48, 64, 80, 96, 128, 160, 192,
        240, 256, 288, 320, 384, 400, 480, 512, 576, 640, 768, 1024)."""
        return bool(_lib.load().tn_plan_has_fft(self.handle))

    def fft_forward(self, image, grid, nimg):
        """image (nimg, *N) -> grid (nimg, *G): apodize, pad and forward DFT.

        3D: the z stage (apodization, placement and DFT along z, into a
        (nimg, Gz, Ny, Nx) workspace), then the (y, x) plane stages on its
        nimg * Gz planes."""
        lib = _lib.load()
        if self.ndim == 3:
            gz, ny, nx = self.grid[0], self.shape[1], self.shape[2]
            with torch.cuda.device(self.device):
                zout = torch.empty((nimg, gz, ny, nx), dtype=torch.complex64, device=self.device)
                _lib.check(lib.tn_fft_z_forward(self.handle, ptr(image), ptr(zout), int(nimg),
                                                stream_ptr(self.device)), "tn_fft_z_forward")
                _lib.check(lib.tn_fft_forward(self.handle, ptr(zout), ptr(grid), int(nimg * gz),
                                              stream_ptr(self.device)), "tn_fft_forward")
            return grid
        _lib.check(lib.tn_fft_forward(self.handle, ptr(image), ptr(grid), int(nimg),
                                      stream_ptr(self.device)), "tn_fft_forward")
        return grid

    def fft_adjoint(self, grid, image, nimg):
        """grid (nimg, *G) -> image (nimg, *N): inverse DFT (1 / prod(G)),
        crop and apodize; ``grid`` is overwritten.

        3D: the (y, x) plane stages on the nimg * Gz grid planes, into a
        (nimg, Gz, Ny, Nx) workspace, then the z stage (inverse DFT along z,
        1 / Gz, crop and apodization)."""
        lib = _lib.load()
        if self.ndim == 3:
            gz, ny, nx = self.grid[0], self.shape[1], self.shape[2]
            with torch.cuda.device(self.device):
                planes = torch.empty((nimg, gz, ny, nx), dtype=torch.complex64, device=self.device)
                _lib.check(lib.tn_fft_adjoint(self.handle, ptr(grid), ptr(planes), int(nimg * gz),
                                              stream_ptr(self.device)), "tn_fft_adjoint")
                _lib.check(lib.tn_fft_z_adjoint(self.handle, ptr(planes), ptr(image), int(nimg),
                                                stream_ptr(self.device)), "tn_fft_z_adjoint")
            return image
        _lib.check(lib.tn_fft_adjoint(self.handle, ptr(grid), ptr(image), int(nimg),
                                      stream_ptr(self.device)), "tn_fft_adjoint")
        return image


def _destroy(handle, device):
    try:
        with torch.cuda.device(device):
            torch.cuda.current_stream(device).synchronize()
            _lib.load().tn_plan_destroy(ctypes.c_void_p(handle))
    except Exception:  # interpreter shutdown
        pass


def check_tensor(t, device, name):
    if t.device != device or t.dtype != torch.complex64 or not t.is_contiguous():
        raise ValueError(f"{name} must be a contiguous complex64 tensor on {device}")
