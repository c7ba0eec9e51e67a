"""Bindings of the solver kernels that replace host loops of small tensor operations."""

import torch

from .. import _lib


def backward_elimination(inverse, b, count, limited=None, limit=0):
    """Removed unknown indices [count] int64 of the one-launch backward elimination;
    inverse [n, n] float64 real, b [n] = inverse q (replacement.py:104-139)."""
    n = len(inverse)
    removed = torch.zeros(count, dtype=torch.int64, device=inverse.device)
    if count == 0:
        return removed
    size = _lib.library().dsplat_elimination_workspace(n, count)
    if size == 0:
        raise ValueError(f"backward elimination kernel supports n <= 4096, count <= n; got {n}, {count}")
    workspace = torch.empty(size // 8 + 1, dtype=torch.float64, device=inverse.device)
    flags = None if limited is None else limited.to(torch.int32).contiguous()
    _lib.call(
        "dsplat_backward_elimination", _lib.ptr(inverse.contiguous()), _lib.ptr(b.contiguous()), n, count,
        _lib.ptr(flags), int(limit), _lib.ptr(removed), _lib.ptr(workspace), size, _lib.stream(inverse.device),
    )  # fmt: skip
    return removed
