"""Linear operators on the NUFFT path (mirror of ``deepmr.linops``' NUFFT / SENSE part)."""

from .base import Add, Compose, Identity, Linop, Multiply  # noqa: F401
from .coil import SenseAdjointOp, SenseOp  # noqa: F401
from .nufft import NUFFTAdjointOp, NUFFTGramOp, NUFFTOp  # noqa: F401
