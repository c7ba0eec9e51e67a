"""Bindings of the surfel kernels: oriented surfels seen as FMCW range profiles.

Mirrors dsplat/kernels/surfel.py of the reference (same functions, arguments and
return shapes). config is an FMCWConfig, positions the sensor positions [A, 3].
weights are per-surfel complex amplitudes, reflectivity times sqrt(area); the area
chain lives in dsplat.surfel.
"""

import torch

from .. import _lib
from . import interleaved


def _check(positions, centers, normals):
    if positions.ndim != 2 or positions.shape[1] != 3:
        raise ValueError(f"positions must be [A, 3], got {tuple(positions.shape)}")
    if centers.ndim != 2 or centers.shape[1] != 3 or normals.shape != centers.shape:
        raise ValueError("centers and normals must both be [S, 3]")
    for t in (positions, centers, normals):
        if t.dtype != torch.float32:
            raise TypeError("surfel kernels take float32 geometry")


def _profiles(config, positions, centers, normals, weights, fresnel_n):
    """Range profiles [A, K, 2] = sum_s weights_s * column_s (kernels/surfel.py:58-108)."""
    _check(positions, centers, normals)
    positions, centers, normals = positions.contiguous(), centers.contiguous(), normals.contiguous()
    weights = weights.contiguous()
    output = centers.new_empty((len(positions), config.num_range_bins, 2))
    cfg = _lib.fmcw(config, fresnel_n)
    _lib.call(
        "dsplat_surfel_render",
        _lib.ctypes.byref(cfg),
        _lib.ptr(positions),
        len(positions),
        _lib.ptr(centers),
        _lib.ptr(normals),
        _lib.ptr(weights),
        len(centers),
        _lib.ptr(output),
        _lib.stream(centers.device),
    )
    return output


def vjp(config, positions, centers, normals, weights, upstream, fresnel_n):
    """Gradients of Re(conj(upstream) . profiles) for centers, normals and complex weights
    (kernels/surfel.py:123-137)."""
    _check(positions, centers, normals)
    positions, centers, normals = positions.contiguous(), centers.contiguous(), normals.contiguous()
    count = len(centers)
    w = interleaved(weights)
    up = interleaved(upstream.reshape(len(positions), config.num_range_bins))
    gc = centers.new_empty((count, 3))
    gn = centers.new_empty((count, 3))
    gw = centers.new_empty((count, 2))
    lib = _lib.library()
    size = lib.dsplat_surfel_vjp_workspace(len(positions), count)
    workspace = centers.new_empty(max(1, size // 4))
    cfg = _lib.fmcw(config, fresnel_n)
    _lib.call(
        "dsplat_surfel_vjp",
        _lib.ctypes.byref(cfg),
        _lib.ptr(positions),
        len(positions),
        _lib.ptr(centers),
        _lib.ptr(normals),
        _lib.ptr(w),
        count,
        _lib.ptr(up),
        _lib.ptr(gc),
        _lib.ptr(gn),
        _lib.ptr(gw),
        _lib.ptr(workspace),
        size,
        _lib.stream(centers.device),
    )
    return gc, gn, torch.view_as_complex(gw)


class _Render(torch.autograd.Function):
    """Forward = render; backward = the VJP recomputed from the saved inputs (the forward
    is never stored), as kernels/surfel.py:167-183."""

    @staticmethod
    def forward(ctx, centers, normals, weights, positions, config, fresnel_n):
        ctx.save_for_backward(centers, normals, weights, positions)
        ctx.config, ctx.fresnel_n = config, fresnel_n
        return _profiles(config, positions, centers, normals, weights, fresnel_n)

    @staticmethod
    def backward(ctx, grad):
        centers, normals, weights, positions = ctx.saved_tensors
        upstream = torch.view_as_complex(grad.contiguous())
        gc, gn, gw = vjp(
            ctx.config, positions, centers, normals, torch.view_as_complex(weights), upstream, ctx.fresnel_n
        )
        return gc, gn, torch.view_as_real(gw), None, None, None


def render(config, positions, centers, normals, weights, fresnel_n):
    """Range profiles [A, K]; differentiable in centers, normals and weights."""
    output = _Render.apply(
        centers.contiguous(), normals.contiguous(), interleaved(weights), positions, config, fresnel_n
    )
    return torch.view_as_complex(output)
