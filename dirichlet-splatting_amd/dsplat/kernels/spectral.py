"""Bindings of the spectral kernels: separable Dirichlet atoms on a spectral grid, in bin units.

Mirrors dsplat/kernels/spectral.py of the reference: design, render (differentiable),
jvp and vjp, with the gradient convention dc = -dk.
"""

import ctypes

import torch

from .. import _lib
from . import interleaved

WINDOWS = ("rect", "hann", "hamming", "kaiser")


def _args(grid, centers):
    """Grid arrays of the C ABI: samples, nfft (host int32), Kaiser taps (device)."""
    dim = len(grid.shape)
    samples = (ctypes.c_int32 * dim)(*grid.samples)
    nfft = (ctypes.c_int32 * dim)(*grid.shape)
    stride = max(grid.samples)
    taps = centers.new_zeros((dim, stride))
    if grid.window == "kaiser":
        for d, n in enumerate(grid.samples):
            w = torch.kaiser_window(n, periodic=False, beta=grid.beta, device=centers.device)
            taps[d, :n] = w / w.sum()
    return dim, samples, nfft, WINDOWS.index(grid.window), float(grid.radius), taps, stride


def design(grid, centers, bins, derivatives=False):
    """Dictionary A [M, S] and, with derivatives, dA/dcenters [M, S, D] (kernels/spectral.py:32-40)."""
    dim, samples, nfft, window, radius, taps, stride = _args(grid, centers)
    a = centers.new_empty((len(bins), len(centers), 2))
    da = centers.new_empty((len(bins), len(centers), dim, 2)) if derivatives else None
    _lib.call(
        "dsplat_spectral_design", dim, samples, nfft, window, radius, _lib.ptr(taps), stride,
        _lib.ptr(centers.float().contiguous()), len(centers), _lib.ptr(bins.float().contiguous()), len(bins),
        _lib.ptr(a), _lib.ptr(da), _lib.stream(centers.device),
    )  # fmt: skip
    a = torch.view_as_complex(a)
    return (a, torch.view_as_complex(da)) if derivatives else a


def _forward(grid, centers, coeff, bins):
    dim, samples, nfft, window, radius, taps, stride = _args(grid, centers)
    out = centers.new_empty((len(bins), 2))
    _lib.call(
        "dsplat_spectral_forward", dim, samples, nfft, window, radius, _lib.ptr(taps), stride,
        _lib.ptr(centers.contiguous()), _lib.ptr(coeff.contiguous()), len(centers),
        _lib.ptr(bins.float().contiguous()), len(bins), _lib.ptr(out), _lib.stream(centers.device),
    )  # fmt: skip
    return out


def _backward(grid, centers, coeff, bins, upstream):
    """(grad centers [S, D], grad coeff [S, 2]) of Re(conj(upstream) . render)."""
    dim, samples, nfft, window, radius, taps, stride = _args(grid, centers)
    count = len(centers)
    out = centers.new_empty((count, dim + 2))
    size = _lib.library().dsplat_spectral_backward_workspace(dim, count, len(bins))
    workspace = centers.new_empty(max(1, size // 4))
    _lib.call(
        "dsplat_spectral_backward", dim, samples, nfft, window, radius, _lib.ptr(taps), stride,
        _lib.ptr(centers.contiguous()), _lib.ptr(coeff.contiguous()), count, _lib.ptr(bins.float().contiguous()),
        len(bins), _lib.ptr(upstream.contiguous()), _lib.ptr(out), _lib.ptr(workspace), size,
        _lib.stream(centers.device),
    )  # fmt: skip
    return out[:, :dim].contiguous(), out[:, dim:].contiguous()


class _Render(torch.autograd.Function):
    """Forward sum of atoms; backward by the deterministic spectral VJP (kernels/spectral.py:168-183)."""

    @staticmethod
    def forward(ctx, centers, coeff, grid, bins):
        ctx.save_for_backward(centers, coeff, bins)
        ctx.grid = grid
        return _forward(grid, centers, coeff, bins)

    @staticmethod
    def backward(ctx, grad):
        centers, coeff, bins = ctx.saved_tensors
        gc, gb = _backward(ctx.grid, centers, coeff, bins, grad.contiguous())
        return gc, gb, None, None


def render(grid, centers, coefficients, bins):
    """Coherent sum of complex atoms at the given bins; differentiable in both inputs."""
    output = _Render.apply(centers.float().contiguous(), interleaved(coefficients), grid, bins)
    return torch.view_as_complex(output)


def jvp(grid, centers, coefficients, delta_centers, delta_coefficients, bins):
    """Directional derivative of the sum at the bins (kernels/spectral.py:192-201)."""
    dim, samples, nfft, window, radius, taps, stride = _args(grid, centers)
    out = centers.new_empty((len(bins), 2))
    _lib.call(
        "dsplat_spectral_jvp", dim, samples, nfft, window, radius, _lib.ptr(taps), stride,
        _lib.ptr(centers.float().contiguous()), _lib.ptr(interleaved(coefficients)),
        _lib.ptr(delta_centers.float().contiguous()), _lib.ptr(interleaved(delta_coefficients)), len(centers),
        _lib.ptr(bins.float().contiguous()), len(bins), _lib.ptr(out), _lib.stream(centers.device),
    )  # fmt: skip
    return torch.view_as_complex(out)


def vjp(grid, centers, coefficients, upstream, bins):
    """Gradients of Re(conj(upstream) . render) for centers and complex coefficients
    (kernels/spectral.py:204-209)."""
    gc, gb = _backward(grid, centers.float().contiguous(), interleaved(coefficients), bins, interleaved(upstream))
    return gc, torch.view_as_complex(gb)
