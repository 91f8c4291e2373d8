"""Iterative solvers used with the NUFFT (mirror of deepmr.optim's CG)."""

from .cg import CGStep, cg_solve  # noqa: F401
