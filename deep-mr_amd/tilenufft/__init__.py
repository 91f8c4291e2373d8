"""tilenufft: MI355X-native Kaiser-Bessel NUFFT with deepmr's plan/apply API.

Modules mirror deepmr's layout for the NUFFT path:
  tilenufft.fft       nufft, nufft_adj, plan_nufft, apply_nufft, apply_nufft_adj
  tilenufft._interp   plan_interpolator, apply_interpolation, apply_gridding
  tilenufft.linops    NUFFTOp, NUFFTAdjointOp, SenseOp, SenseAdjointOp, algebra
  tilenufft.optim     CGStep, cg_solve
  tilenufft.calib     find_acs (non-Cartesian calibration through nufft_adj)
  tilenufft.dcf       voronoi, analytical_dcf, angular_compensation
  tilenufft.sampling  radial, radial_proj, spiral (trajectory generators)
  tilenufft.sense     CG-SENSE state and the fixed-order SENSE / CG kernels
  tilenufft.serve     continuous-batching CG-SENSE server
The compute path is the native library libtilenufft.so (HIP, gfx950).
"""

from . import _interp, _signal, autotune, batch, calib, dcf, fft, linops, optim, recon, sampling, sense, serve, testing  # noqa: F401
from ._lib import version  # noqa: F401
from ._signal import resize  # noqa: F401

__version__ = "0.1.0"
