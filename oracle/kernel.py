"""Unit-peak, negative-exponent Dirichlet kernel in float64 (oracle).

D(d) = (1/S) sum_{n<S} exp(-2 pi i n d / N): the range response of S ADC samples
zero-padded to an N-point FFT, evaluated at an offset of d bins.

Follows `dsplat/kernels/common.slang:10-45` and `tests/reference/kernel.py:8-57`.
Pinned by `dirichlet_bruteforce` (the defining DFT sum itself, a textbook
geometric series) in tests/test_oracle.py, including d near 0 (Taylor branch),
the zeros d = m N / S, the periodicity N and the derivative.
"""

import math

import numpy as np


def dirichlet(delta, samples, n_fft=None):
    """(D, dD/ddelta) in closed form with the reference's Taylor branch.

    `common.slang:10-33`: x = delta wrapped to [-N/2, N/2]; ratio =
    sin(a x) / (S sin(b x)), a = pi S / N, b = pi / N; |x| < 1e-3 uses
    1 + c2 x^2 + c4 x^4; the phase is exp(-i pi (S - 1) x / N).
    """
    n_fft = samples if n_fft is None else n_fft
    delta = np.asarray(delta, dtype=np.float64)
    x = delta - n_fft * np.round(delta / n_fft)
    a, b = math.pi * samples / n_fft, math.pi / n_fft
    h = math.pi * (samples - 1) / n_fft
    near = np.abs(x) < 1e-3
    safe = np.where(near, 1.0, x)
    sa, sb = np.sin(a * safe), np.sin(b * safe)
    ratio = sa / (samples * sb)
    deriv = (a * np.cos(a * safe) * sb - b * sa * np.cos(b * safe)) / (samples * sb * sb)
    c2 = (b * b - a * a) / 6
    c4 = (3 * a**4 + 7 * b**4 - 10 * a * a * b * b) / 360
    ratio = np.where(near, 1 + c2 * x * x + c4 * x**4, ratio)
    deriv = np.where(near, 2 * c2 * x + 4 * c4 * x**3, deriv)
    phase = np.exp(-1j * h * x)
    return phase * ratio, phase * (deriv - 1j * h * ratio)


def dirichlet_bruteforce(delta, samples, n_fft=None, taps=None):
    """The defining sum: D = sum_n t_n exp(-2 pi i n d / N), t_n = 1/S unless taps given.

    Also returns the exact derivative sum_n t_n (-2 pi i n / N) exp(...).
    """
    n_fft = samples if n_fft is None else n_fft
    delta = np.asarray(delta, dtype=np.float64)
    n = np.arange(samples, dtype=np.float64)
    t = np.full(samples, 1.0 / samples) if taps is None else np.asarray(taps, np.float64)
    omega = -2j * math.pi * n / n_fft
    e = np.exp(delta[..., None] * omega)
    return (e * t).sum(-1), (e * t * omega).sum(-1)


WINDOW_WEIGHT = {"hann": 0.5, "hamming": 0.46 / 1.08}


def window_taps(window, samples, beta=14.0):
    """Normalized taps of each window: rect 1/S; periodic Hann/Hamming; symmetric Kaiser.

    Hann/Hamming as the 3-term shifted Dirichlet sum of `common.slang:35-45` equal a
    DFT with taps (1 - 2 w cos(2 pi n / S)) / S; Kaiser taps follow
    `kernels/spectral.py:11-15` (torch.kaiser_window(periodic=False) / sum).
    """
    n = np.arange(samples, dtype=np.float64)
    if window == "rect":
        return np.full(samples, 1.0 / samples)
    if window in WINDOW_WEIGHT:
        return (1 - 2 * WINDOW_WEIGHT[window] * np.cos(2 * math.pi * n / samples)) / samples
    if window == "kaiser":
        if samples == 1:
            return np.ones(1)
        arg = beta * np.sqrt(np.clip(1 - (2 * n / (samples - 1) - 1) ** 2, 0, None))
        w = np.i0(arg) / np.i0(beta)
        return w / w.sum()
    raise ValueError(window)


def windowed(delta, samples, n_fft=None, window="rect", beta=14.0):
    """Windowed kernel and derivative (`common.slang:35-45`, `spectral.slang:15-23`).

    rect/hann/hamming use the closed form; kaiser the explicit tap sum.
    """
    n_fft = samples if n_fft is None else n_fft
    if window == "rect":
        return dirichlet(delta, samples, n_fft)
    if window == "kaiser":
        delta = np.asarray(delta, np.float64)
        delta = delta - n_fft * np.round(delta / n_fft)
        return dirichlet_bruteforce(delta, samples, n_fft, window_taps("kaiser", samples, beta))
    w = WINDOW_WEIGHT[window]
    shift = n_fft / samples
    k, dk = dirichlet(delta, samples, n_fft)
    km, dm = dirichlet(np.asarray(delta) - shift, samples, n_fft)
    kp, dp = dirichlet(np.asarray(delta) + shift, samples, n_fft)
    return k - w * (km + kp), dk - w * (dm + dp)
