"""Dense float64 products and responses (oracle).

normal_equations is A^H A / M and A^H y / M of a dense design (`dsplat/varpro.py:91-120`;
the reference tests compare its streamed Gram with this, `tests/test_solver.py:30-43`);
range_response restates `dsplat/kernels/response.slang:2-26`; energy is sum |x - y|^2.
"""

import numpy as np

from .kernel import dirichlet


def normal_equations(a, y):
    a = np.asarray(a, np.complex128)
    y = np.asarray(y, np.complex128)
    m = len(a)
    return a.conj().T @ a / m, a.conj().T @ y / m


def energy(x, y):
    return float((np.abs(np.asarray(x, np.complex128) - np.asarray(y, np.complex128)) ** 2).sum())


def range_response(low, high, count, bins, samples, n_fft):
    centers = low + (high - low) * np.arange(count) / max(count - 1, 1)
    delta = np.arange(bins)[:, None] - centers[None, :]
    return dirichlet(delta, samples, n_fft)[0]
