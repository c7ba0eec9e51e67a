"""Oriented surfels seen by an FMCW sensor, in float64 (oracle).

Column model (`dsplat/surfel.py:1-33`, `kernels/scattering.slang:1-50`,
`tests/reference/oracle.py:34-41`): for sensor p, surfel center c and unit normal n,

    r = |p - c| + 1e-12,  u = (p - c) / r,  cos = u . n
    amp = cos^2 F(cos) lambda / (8 pi r)           (0 when cos <= 0)
    rot = exp(2 pi i (fc tau + K tau (t0 - tau / 2))),  tau = 2 r / c0
    column[a, k, s] = amp rot D(k - r / spacing)

render = columns @ weights; JVP/VJP are its exact derivatives; design multiplies by
sqrt(area). The `synthesize` ADC oracle (`dsplat/groundtruth.py:40-90`) never
evaluates D and pins `columns` (tests/test_oracle.py); finite differences pin the
Jacobian, the adjoint identity pins the VJP against the JVP.

Inputs are float32 arrays (as the GPU sees them), promoted to float64 here.
"""

import math

import numpy as np

from .kernel import dirichlet


class Chirp:
    """Derived constants of an FMCWConfig-like object (`dsplat/fmcw.py:29-65`)."""

    def __init__(self, cfg):
        self.fc = float(cfg.fc_hz)
        self.bandwidth = float(cfg.bandwidth_hz)
        self.samples = int(cfg.adc_samples)
        self.sweep = float(cfg.sweep_time_s)
        self.t0 = float(cfg.adc_start_time_s)
        self.pad = int(cfg.pad_factor)
        self.c0 = float(cfg.c0_m_s)
        self.lam = self.c0 / self.fc
        self.slope = self.bandwidth / self.sweep
        self.rate = self.samples / self.sweep
        self.n_fft = self.samples * self.pad
        self.bins = self.n_fft // 2
        self.spacing = self.c0 / (2 * self.bandwidth * self.pad)


def _f64(x):
    return np.asarray(x, dtype=np.float64)


def fresnel(cosine, n):
    """F(cos) and dF/dcos of the s-polarized dielectric (`scattering.slang:16-25`)."""
    if n <= 1:
        return np.ones_like(cosine), np.zeros_like(cosine)
    root = np.sqrt(n * n - 1 + cosine * cosine)
    return (root - cosine) / (root + cosine), -2 * (root - cosine) / (root * (root + cosine))


def scattering(cosine, r, fresnel_n, lam):
    """amp = cos^2 F lambda / (8 pi r) and d amp / d cos; 0 for back faces (`scattering.slang:27-38`)."""
    f, df = fresnel(cosine, fresnel_n)
    path = lam / (8 * math.pi * r)
    visible = cosine > 0
    amp = np.where(visible, path * cosine * cosine * f, 0.0)
    amp_cos = np.where(visible, path * (2 * cosine * f + cosine * cosine * df), 0.0)
    return amp, amp_cos


def geometry(positions, centers, normals, chirp, fresnel_n=0.0):
    """Per pair [A, S]: r, u [A,S,3], cos, amp, amp_cos, rot, phase rate, k0.

    `surfel.slang:63-85`, `scattering.slang:40-50` in float64.
    """
    d = _f64(positions)[:, None, :] - _f64(centers)[None, :, :]
    r = np.sqrt((d * d).sum(-1)) + 1e-12
    u = d / r[..., None]
    cosine = (u * _f64(normals)[None]).sum(-1)
    amp, amp_cos = scattering(cosine, r, fresnel_n, chirp.lam)
    tau = 2 * r / chirp.c0
    cycles = chirp.fc * tau + chirp.slope * tau * (chirp.t0 - 0.5 * tau)
    cycles = cycles - np.round(cycles)
    rot = np.exp(2j * math.pi * cycles)
    rate = 2 * math.pi * 2 / chirp.c0 * (chirp.fc + chirp.slope * (chirp.t0 - 2 * r / chirp.c0))
    k0 = r / chirp.spacing
    return dict(r=r, u=u, cos=cosine, amp=amp, amp_cos=amp_cos, rot=rot, rate=rate, k0=k0)


def _kernel(geo, chirp):
    """D and D' [A, K, S] at every observed bin."""
    k = np.arange(chirp.bins, dtype=np.float64)
    delta = k[None, :, None] - geo["k0"][:, None, :]
    return dirichlet(delta, chirp.samples, chirp.n_fft)


def columns(positions, centers, normals, cfg, fresnel_n=0.0):
    """Columns [A, K, S] complex128 without area (`tests/reference/oracle.py:34-41`)."""
    chirp = Chirp(cfg)
    geo = geometry(positions, centers, normals, chirp, fresnel_n)
    value, _ = _kernel(geo, chirp)
    return value * (geo["amp"] * geo["rot"])[:, None, :]


def jacobian(positions, centers, normals, cfg, fresnel_n=0.0):
    """Columns [A,K,S] and their center and normal derivatives [A,K,S,3].

    d col / d c_j = rot [ (amp u_j / r + amp_cos (cos u_j - n_j) / r) D
                          - i amp rate u_j D + amp k0pm u_j D' ],
    d col / d n_j = rot amp_cos u_j D   (`correlate.slang:239-281`).
    """
    chirp = Chirp(cfg)
    geo = geometry(positions, centers, normals, chirp, fresnel_n)
    value, slope = _kernel(geo, chirp)
    amp, amp_cos, r, u, cos, rot = (geo[k] for k in ("amp", "amp_cos", "r", "u", "cos", "rot"))
    n = _f64(normals)[None]
    da = amp[..., None] * u / r[..., None] + amp_cos[..., None] * (cos[..., None] * u - n) / r[..., None]
    k0pm = 1 / chirp.spacing
    col = value * (amp * rot)[:, None, :]
    dcen = rot[:, None, :, None] * (
        da[:, None] * value[..., None]
        - 1j * (amp * geo["rate"])[:, None, :, None] * u[:, None] * value[..., None]
        + (amp * k0pm)[:, None, :, None] * u[:, None] * slope[..., None]
    )
    dnrm = (rot * amp_cos)[:, None, :, None] * u[:, None] * value[..., None]
    return col, dcen, dnrm


def _chunks(count, size):
    for start in range(0, count, size):
        yield slice(start, min(count, start + size))


def render(positions, centers, normals, weights, cfg, fresnel_n=0.0, chunk=64):
    """Range profiles [A, K] = sum_s w_s column_s (`kernels/surfel.py:58-108`)."""
    w = np.asarray(weights, np.complex128)
    out = []
    for part in _chunks(len(positions), chunk):
        out.append(columns(positions[part], centers, normals, cfg, fresnel_n) @ w)
    return np.concatenate(out)


def jvp(positions, centers, normals, weights, dcenters, dnormals, dweights, cfg, fresnel_n=0.0, chunk=8):
    """Directional derivative [A, K] (`surfel.slang:146-221`)."""
    w = np.asarray(weights, np.complex128)
    dw = np.asarray(dweights, np.complex128)
    dc, dn = _f64(dcenters), _f64(dnormals)
    out = []
    for part in _chunks(len(positions), chunk):
        col, dcen, dnrm = jacobian(positions[part], centers, normals, cfg, fresnel_n)
        tangent = (dcen * dc[None, None]).sum(-1) + (dnrm * dn[None, None]).sum(-1)
        out.append(tangent @ w + col @ dw)
    return np.concatenate(out)


def vjp(positions, centers, normals, weights, upstream, cfg, fresnel_n=0.0, chunk=8):
    """Gradients of Re(conj(upstream) . render): gc [S,3], gn [S,3], gw [S] complex.

    gw = A^H u; gc_j = Re sum conj(u) w d col / d c_j (`surfel.slang:227-311`).
    """
    w = np.asarray(weights, np.complex128)
    up = np.asarray(upstream, np.complex128)
    count = len(centers)
    gc, gn = np.zeros((count, 3)), np.zeros((count, 3))
    gw = np.zeros(count, np.complex128)
    for part in _chunks(len(positions), chunk):
        col, dcen, dnrm = jacobian(positions[part], centers, normals, cfg, fresnel_n)
        u = up[part]
        gw += np.einsum("aks,ak->s", col.conj(), u)
        gc += (np.einsum("ak,aksj->sj", u.conj(), dcen) * w[:, None]).real
        gn += (np.einsum("ak,aksj->sj", u.conj(), dnrm) * w[:, None]).real
    return gc, gn, gw


def design(positions, centers, normals, areas, cfg, fresnel_n=0.0, real=False):
    """Dictionary rows: complex [A K, S], or real [2 K A, S] with row (2k+part) A + a.

    `surfel.slang:313-378`, `kernels/surfel.py:199-220`.
    """
    col = columns(positions, centers, normals, cfg, fresnel_n) * np.sqrt(_f64(areas))[None, None]
    count_a, bins, count_s = col.shape
    if not real:
        return col.reshape(count_a * bins, count_s)
    out = np.empty((bins, 2, count_a, count_s))
    out[:, 0] = col.real.transpose(1, 0, 2)
    out[:, 1] = col.imag.transpose(1, 0, 2)
    return out.reshape(2 * bins * count_a, count_s)


def correlate(positions, centers, normals, residual, cfg, fresnel_n=0.0, chunk=32):
    """a^H r [S] and |a|^2 [S] per candidate, without area (`correlate.slang:78-108`)."""
    res = np.asarray(residual, np.complex128)
    corr = np.zeros(len(centers), np.complex128)
    norm = np.zeros(len(centers))
    for part in _chunks(len(positions), chunk):
        col = columns(positions[part], centers, normals, cfg, fresnel_n)
        corr += np.einsum("aks,ak->s", col.conj(), res[part])
        norm += (np.abs(col) ** 2).sum((0, 1))
    return corr, norm


def correlate_gradient(positions, centers, normals, residual, cfg, fresnel_n=0.0, include_normals=False, chunk=16):
    """corr, norm and d corr, d norm by centers (and normals) (`correlate.slang:174-235`).

    d(a^H r) = da^H r; d|a|^2 = 2 Re(da^H a).
    """
    res = np.asarray(residual, np.complex128)
    width = 6 if include_normals else 3
    count = len(centers)
    corr = np.zeros(count, np.complex128)
    norm = np.zeros(count)
    dcorr = np.zeros((count, width), np.complex128)
    dnorm = np.zeros((count, width))
    for part in _chunks(len(positions), chunk):
        col, dcen, dnrm = jacobian(positions[part], centers, normals, cfg, fresnel_n)
        da = np.concatenate((dcen, dnrm), -1)[..., :width]
        corr += np.einsum("aks,ak->s", col.conj(), res[part])
        norm += (np.abs(col) ** 2).sum((0, 1))
        dcorr += np.einsum("aksj,ak->sj", da.conj(), res[part])
        dnorm += 2 * np.einsum("aksj,aks->sj", da.conj(), col).real
    return corr, norm, dcorr, dnorm


def correlate_orientations(positions, centers, normals, bank, residual, cfg, fresnel_n=0.0):
    """corr, norm [S, 1 + B] for each surfel's own normal and every bank normal.

    `correlate.slang:112-169` (exact per-bin path; column 0 is the own normal).
    """
    count, extra = len(centers), len(bank)
    corr = np.zeros((count, 1 + extra), np.complex128)
    norm = np.zeros((count, 1 + extra))
    corr[:, 0], norm[:, 0] = correlate(positions, centers, normals, residual, cfg, fresnel_n)
    for j in range(extra):
        rep = np.broadcast_to(_f64(bank[j]), (count, 3))
        corr[:, 1 + j], norm[:, 1 + j] = correlate(positions, centers, rep, residual, cfg, fresnel_n)
    return corr, norm


def certificate(corr, norm, residual, mode):
    """|s|^2 / (|a|^2 |r|^2), s = Re(a^H r) (real) or a^H r (`certificate.py:11-31`)."""
    corr = np.asarray(corr, np.complex128)
    norm = np.clip(np.asarray(norm, np.float64), 1e-60, None)
    power = max(float((np.abs(np.asarray(residual, np.complex128)) ** 2).sum()), 1e-30)
    s = corr.real if mode == "real" else corr
    return np.abs(s) ** 2 / norm / power


def planar_array(side, pitch, standoff=0.35):
    """side x side positions [side^2, 3] in the plane z = 0, row-major (`groundtruth.py:16-37`)."""
    t = (np.arange(side) - (side - 1) / 2) * pitch
    x, y = np.meshgrid(t, t, indexing="ij")
    return np.stack([x.ravel(), y.ravel(), np.zeros(side * side)], -1).astype(np.float32)


def synthesize(positions, centers, normals, amplitudes, cfg, fresnel_n=0.0):
    """FP64 ADC beat-tone synthesis plus range FFT / S (`groundtruth.py:58-90`).

    The beat frequency 2 K r / c0 sampled at t_n = n / rate, summed over surfels in
    complex128, FFT to n_fft and cut to the positive bins. Never evaluates D.
    """
    chirp = Chirp(cfg)
    time = np.arange(chirp.samples, dtype=np.float64) / chirp.rate
    amps = np.asarray(amplitudes, np.complex128)
    out = []
    for part in _chunks(len(positions), 64):
        d = _f64(positions[part])[:, None, :] - _f64(centers)[None]
        distance = np.maximum(np.sqrt((d * d).sum(-1)), 1e-9)
        cosine = np.clip((d / distance[..., None] * _f64(normals)[None]).sum(-1), 0, None)
        amplitude = cosine**2 * chirp.lam / (8 * math.pi * distance)
        if fresnel_n > 1:
            root = np.sqrt(fresnel_n**2 - 1 + cosine**2)
            amplitude = amplitude * (root - cosine) / (root + cosine)
        tau = 2 * distance / chirp.c0
        phase = 2 * math.pi * (chirp.fc * tau + chirp.slope * tau * (chirp.t0 - 0.5 * tau))
        echo = amplitude * np.exp(1j * phase) * amps[None]
        beat = 2 * chirp.slope * distance / chirp.c0
        adc = np.einsum("as,ast->at", echo, np.exp(2j * math.pi * beat[..., None] * time))
        spectrum = np.fft.fft(adc, n=chirp.n_fft)[:, : chirp.bins]
        out.append(spectrum / chirp.samples)
    return np.concatenate(out)
