"""Linear-operator algebra used by the NUFFT callers (CG-SENSE).

Same behaviour as deepmr's Linop family (src/deepmr/linops/base.py:10-247):
operators are ``nn.Module``s with an ``ndim`` (number of spatial axes), an
adjoint through ``.H``, and ``+``, ``*`` (composition / scaling) and ``-``.
"""

__all__ = ["Linop", "Add", "Compose", "Multiply", "Identity"]

import numpy as np
import torch
import torch.nn as nn


class Linop(nn.Module):
    """Base class: subclasses implement ``forward`` and ``_adjoint_linop``."""

    def __init__(self, ndim):
        super().__init__()
        self.ndim = ndim

    def _adjoint_linop(self):
        raise NotImplementedError

    @property
    def H(self):
        """Adjoint operator."""
        return self._adjoint_linop()

    def __add__(self, other):
        return Add([self, other])

    def __mul__(self, other):
        if np.isscalar(other) or (isinstance(other, torch.Tensor) and other.dim() == 0):
            return Multiply(self, other)
        if isinstance(other, Linop):
            return Compose([self, other])
        if isinstance(other, torch.Tensor):
            return self(other)
        raise NotImplementedError(f"cannot multiply a Linop by {type(other)}")

    def __rmul__(self, other):
        if np.isscalar(other) or (isinstance(other, torch.Tensor) and other.dim() == 0):
            return Multiply(self, other)
        return NotImplemented

    def __sub__(self, other):
        return self + (-other)

    def __neg__(self):
        return -1 * self

    def to(self, *args, **kwargs):
        for key, val in list(self.__dict__.items()):
            if isinstance(val, torch.Tensor):
                self.__dict__[key] = val.to(*args, **kwargs)
        return self


def _same_ndim(linops):
    dims = {op.ndim for op in linops}
    if len(dims) != 1:
        raise ValueError("all operators must have the same spatial dimensionality")
    return dims.pop()


class Add(Linop):
    """(A + B) x = A x + B x."""

    def __init__(self, linops):
        super().__init__(_same_ndim(linops))
        self.linops = list(linops)

    def forward(self, x):
        out = None
        for op in self.linops:
            y = op(x)
            out = y if out is None else out + y
        return out

    def _adjoint_linop(self):
        return Add([op.H for op in self.linops])


class Compose(Linop):
    """(A * B) x = A(B(x)); the adjoint reverses the order."""

    def __init__(self, linops):
        super().__init__(_same_ndim(linops))
        self.linops = list(linops)

    def forward(self, x):
        for op in reversed(self.linops):
            x = op(x)
        return x

    def _adjoint_linop(self):
        return Compose([op.H for op in reversed(self.linops)])


class Multiply(Linop):
    """(a * A) x = a * A(x)."""

    def __init__(self, linop, a):
        super().__init__(linop.ndim)
        self.linop = linop
        self.a = a

    def forward(self, x):
        return self.a * self.linop(x)

    def _adjoint_linop(self):
        a = self.a.conj() if isinstance(self.a, torch.Tensor) else np.conj(self.a)
        return Multiply(self.linop.H, a)


class Identity(Linop):
    """I x = x."""

    def forward(self, x):
        return x

    def _adjoint_linop(self):
        return self
