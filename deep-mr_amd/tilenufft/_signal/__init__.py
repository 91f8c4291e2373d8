"""Signal helpers on the NUFFT path."""

from .resize import resize  # noqa: F401
