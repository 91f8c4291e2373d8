"""Centred zero-pad / crop (drop-in for src/deepmr/_signal/resize.py:11-102).

The NUFFT itself never calls this: its pad and crop are fused into the
native apodization kernels. Kept for API parity and as a torch utility.
"""

__all__ = ["resize"]

import numpy as np
import torch


def resize(input, oshape):
    """Zero-pad or crop the trailing axes to ``oshape`` around the centre N//2."""
    isnumpy = isinstance(input, np.ndarray)
    x = torch.as_tensor(input)
    oshape = [oshape] if isinstance(oshape, int) else list(oshape)
    ishape = list(x.shape)
    nd = max(len(ishape), len(oshape))
    full_i = [1] * (nd - len(ishape)) + ishape
    full_o = [1] * (nd - len(oshape)) + oshape
    lead = [max(a, b) for a, b in zip(full_i, full_o)]
    ishape1 = lead[: nd - len(ishape)] + ishape
    oshape1 = lead[: nd - len(oshape)] + oshape
    if ishape1 == oshape1:
        return input
    src = [max(i // 2 - o // 2, 0) for i, o in zip(ishape1, oshape1)]
    dst = [max(o // 2 - i // 2, 0) for i, o in zip(ishape1, oshape1)]
    n = [min(i - s, o - d) for i, s, o, d in zip(ishape1, src, oshape1, dst)]
    out = torch.zeros(oshape1, dtype=x.dtype, device=x.device)
    out[tuple(slice(d, d + c) for d, c in zip(dst, n))] = x.reshape(ishape1)[
        tuple(slice(s, s + c) for s, c in zip(src, n))
    ]
    return out.numpy(force=True) if isnumpy else out
