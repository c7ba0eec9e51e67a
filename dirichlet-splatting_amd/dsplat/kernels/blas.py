"""Dense products on MFMA: the symmetric rank-k update the reference binds from cuBLAS
(dsplat/kernels/cublas.py) and the fused split-precision Gram."""

import torch

from .. import _lib


def syrk(rows, out):
    """Add rows^T rows to the upper triangle of out [n, n]; the lower triangle is untouched
    (kernels/cublas.py:31-44). rows [m, n] float32, out float32 contiguous."""
    if rows.dtype != torch.float32 or out.dtype != torch.float32:
        raise TypeError("syrk takes float32 tensors")
    if not out.is_contiguous() or out.shape != (rows.shape[1], rows.shape[1]):
        raise ValueError("out must be a contiguous [n, n] tensor")
    rows = rows.contiguous()
    _lib.call("dsplat_syrk", _lib.ptr(rows), rows.shape[0], rows.shape[1], _lib.ptr(out), _lib.stream(rows.device))
    return out
