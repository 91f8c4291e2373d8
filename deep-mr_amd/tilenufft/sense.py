"""CG-SENSE state and the native SENSE / CG kernels shared by the server and
the plain path.

E = F * C with F the NUFFT and C the coil maps (src/deepmr/recon/alg/
linop.py:108-148); CG runs on E^H E x = E^H y (src/deepmr/optim/cg.py:
150-212). ``CGState`` holds one reconstruction's device vectors; a
``SenseGroup`` binds several reconstructions of one image size to the native
expand / combine / CG-update launches (``tn_sense_expand``,
``tn_sense_combine``, ``tn_cg_step``). Their reductions run in a fixed order
that depends only on the image size, so one reconstruction gives the same
bits alone or in a group, and through the server or the plain path.
"""

__all__ = ["CGState", "SenseGroup", "START", "ITERATE", "IDLE"]

import ctypes

import numpy as np
import torch

from . import _core, _lib

START, ITERATE, IDLE = 0, 1, 2


class SenseJob(ctypes.Structure):
    """Mirror of ``tn_sense_job``."""

    _fields_ = [
        ("smaps", ctypes.c_void_p),
        ("x", ctypes.c_void_p),
        ("r", ctypes.c_void_p),
        ("p", ctypes.c_void_p),
        ("ap", ctypes.c_void_p),
        ("scal", ctypes.c_void_p),
        ("img_off", ctypes.c_int64),
        ("ncoils", ctypes.c_int32),
        ("mode", ctypes.c_int32),
    ]


class CGState:
    """Device-resident CG-SENSE vectors of one 2D reconstruction.

    smaps: (ncoils, n, n) complex64 on the device. ``x`` is the current
    iterate; ``it`` counts the CG iterations done (-1 before the start).
    """

    def __init__(self, smaps, device):
        self.smaps = smaps.to(device, torch.complex64).contiguous()
        self.ncoils = int(self.smaps.shape[0])
        self.shape = tuple(int(s) for s in self.smaps.shape[1:])
        self.npix = int(np.prod(self.shape))
        buf = torch.zeros((4, *self.shape), dtype=torch.complex64, device=device)
        self.x, self.r, self.p, self.ap = buf[0], buf[1], buf[2], buf[3]
        self.scal = torch.zeros(2, dtype=torch.float32, device=device)
        self.it = -1

    def nbytes(self):
        return 8 * (self.ncoils + 4) * self.npix + 8


class SenseGroup:
    """Reconstructions of one image size bound to the native SENSE / CG launches.

    ``img_off[k]`` is the offset (complex elements) of reconstruction k's first
    coil image in the coil-image arena passed to ``expand`` / ``combine``.
    """

    def __init__(self, states, img_off, device):
        if not states:
            raise ValueError("empty group")
        npix = {s.npix for s in states}
        if len(npix) != 1:
            raise ValueError("a group shares one image size")
        self.states = list(states)
        self.npix = npix.pop()
        self.img_off = [int(o) for o in img_off]
        self.device = device
        self.work = torch.empty(max(1, int(_lib.load().tn_cg_work_floats(len(states), self.npix))),
                                dtype=torch.float32, device=device)
        self._table = None
        self._modes = None

    def set_modes(self, modes):
        """Upload the records with one mode per reconstruction (START,
        ITERATE or IDLE); unchanged modes reuse the uploaded table."""
        modes = tuple(int(m) for m in modes)
        if modes == self._modes:
            return
        recs = (SenseJob * len(self.states))()
        for rec, s, off, m in zip(recs, self.states, self.img_off, modes):
            rec.smaps = s.smaps.data_ptr()
            rec.x, rec.r, rec.p, rec.ap = s.x.data_ptr(), s.r.data_ptr(), s.p.data_ptr(), s.ap.data_ptr()
            rec.scal = s.scal.data_ptr()
            rec.img_off = off
            rec.ncoils = s.ncoils
            rec.mode = m
        host = torch.frombuffer(bytearray(bytes(recs)), dtype=torch.uint8)
        if torch.cuda.is_available():
            host = host.pin_memory()
        # the caching host allocator keeps the pinned copy alive until the
        # asynchronous upload has run
        self._table = host.to(self.device, non_blocking=True)
        self._modes = modes

    def _jobs(self):
        if self._table is None:
            raise RuntimeError("set_modes() first")
        return ctypes.c_void_p(self._table.data_ptr())

    def expand(self, img):
        """img[off_k + c] = smaps_k[c] * p_k for reconstructions in ITERATE mode."""
        _lib.check(_lib.load().tn_sense_expand(self._jobs(), len(self.states), self.npix, _core.ptr(img),
                                               _core.stream_ptr(self.device)), "tn_sense_expand")

    def combine(self, img):
        """ap_k = sum_c conj(smaps_k[c]) * img[off_k + c] (START / ITERATE)."""
        _lib.check(_lib.load().tn_sense_combine(self._jobs(), len(self.states), self.npix, _core.ptr(img),
                                                _core.stream_ptr(self.device)), "tn_sense_combine")

    def cg_step(self):
        """One CG update per reconstruction from its ap (START: x = 0,
        r = p = ap)."""
        _lib.check(_lib.load().tn_cg_step(self._jobs(), len(self.states), self.npix, _core.ptr(self.work),
                                          _core.stream_ptr(self.device)), "tn_cg_step")
