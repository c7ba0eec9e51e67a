"""Separable windowed Dirichlet atoms on a spectral grid in float64 (oracle).

Follows `dsplat/kernels/spectral.slang:4-30` and `tests/reference/oracle.py:10-31`:
the atom at bin m of a center c is prod_d k_d(m_d - c_d), each axis the windowed
kernel of oracle.kernel (itself pinned against tapped DFT sums), zeroed outside
the radius; derivatives are -dk_d times the other axes. Pinned in tests/test_oracle.py
by a brute-force 2-D DFT of a windowed tone.
"""

import numpy as np

from .kernel import windowed


def design(shape, samples, window, beta, radius, centers, bins):
    """A [M, S] complex128 and dA/dcenters [M, S, D]."""
    centers = np.asarray(centers, np.float64)
    bins = np.asarray(bins, np.float64)
    parts = []
    for d, (n, f) in enumerate(zip(samples, shape, strict=True)):
        delta = bins[:, None, d] - centers[None, :, d]
        k, dk = windowed(delta, n, f, window, beta)
        if radius:
            inside = np.abs(delta - f * np.round(delta / f)) <= radius
            k, dk = k * inside, dk * inside
        parts.append((k, dk))
    a = np.ones((len(bins), len(centers)), np.complex128)
    for k, _ in parts:
        a = a * k
    derivatives = []
    for d, (_, dk) in enumerate(parts):
        value = -dk
        for other, (k, _) in enumerate(parts):
            if other != d:
                value = value * k
        derivatives.append(value)
    return a, np.stack(derivatives, -1)


def grid_bins(shape):
    axes = [np.arange(n, dtype=np.float64) for n in shape]
    return np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, len(shape))
