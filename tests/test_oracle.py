"""Pins the float64 oracle itself: brute force, physics, special cases, finite differences.

None of these retypes an oracle formula: each compares it with an independent
computation (the defining DFT sum, the ADC beat-signal chain, closed-form special
cases, central differences, the adjoint identity).
"""

import math
import types

import numpy as np
import pytest
from conftest import aperture_grid, close, rel_l2, surfels

from oracle import kernel as K
from oracle import surfel as O


def config(pad=2, samples=128):
    return types.SimpleNamespace(
        fc_hz=120e9, bandwidth_hz=15e9, adc_samples=samples, sweep_time_s=1e-5,
        adc_start_time_s=0.0, pad_factor=pad, c0_m_s=299792458.0,
    )  # fmt: skip


@pytest.mark.parametrize("samples,n_fft", [(128, 256), (128, 512), (16, 16), (7, 21)])
def test_dirichlet_equals_the_defining_dft_sum(samples, n_fft):
    delta = np.concatenate([np.linspace(-2 * n_fft, 2 * n_fft, 4001), [0.0, 1e-5, -4e-4, 9.99e-4, 1.2e-3, 0.5]])
    value, deriv = K.dirichlet(delta, samples, n_fft)
    exact, exact_deriv = K.dirichlet_bruteforce(delta, samples, n_fft)
    assert np.abs(value - exact).max() < 1e-12
    assert np.abs(deriv - exact_deriv).max() < 1e-9 * np.abs(exact_deriv).max()


def test_dirichlet_special_values():
    # unit peak, zeros at d = m N / S, period N
    value, _ = K.dirichlet(np.array([0.0, 2.0, 4.0, -6.0, 256.0, 258.0]), 128, 256)
    np.testing.assert_allclose(value, [1, 0, 0, 0, 1, 0], atol=1e-14)


@pytest.mark.parametrize("window", ["hann", "hamming", "kaiser"])
def test_windowed_kernels_equal_tapped_dft_sums(window):
    delta = np.linspace(-40, 40, 801) + 0.013
    value, deriv = K.windowed(delta, 24, 48, window, beta=6.0)
    exact, exact_deriv = K.dirichlet_bruteforce(delta, 24, 48, K.window_taps(window, 24, 6.0))
    assert np.abs(value - exact).max() < 1e-12
    assert np.abs(deriv - exact_deriv).max() < 1e-10


def test_window_taps_are_the_textbook_windows():
    n = np.arange(16)
    hann = 0.5 - 0.5 * np.cos(2 * math.pi * n / 16)  # periodic Hann
    np.testing.assert_allclose(K.window_taps("hann", 16), hann / hann.sum(), atol=1e-15)
    hamming = 0.54 - 0.46 * np.cos(2 * math.pi * n / 16)
    np.testing.assert_allclose(K.window_taps("hamming", 16), hamming / hamming.sum(), atol=1e-15)
    kaiser = np.kaiser(16, 6.0)
    np.testing.assert_allclose(K.window_taps("kaiser", 16, 6.0), kaiser / kaiser.sum(), atol=1e-15)


def test_scattering_special_cases():
    cos = np.array([-0.5, 0.0, 0.3, 1.0])
    amp, _ = O.scattering(cos, np.full(4, 0.3), 0.0, 0.0025)
    path = 0.0025 / (8 * math.pi * 0.3)
    np.testing.assert_allclose(amp, [0, 0, 0.09 * path, path], rtol=1e-15)
    # normal incidence on a dielectric: F = (n - 1) / (n + 1)
    amp16, _ = O.scattering(np.array([1.0]), np.array([0.3]), 1.6, 0.0025)
    np.testing.assert_allclose(amp16, path * 0.6 / 2.6, rtol=1e-14)
    # d amp / d cos by central differences
    c = np.linspace(0.05, 0.95, 19)
    for n in (0.0, 1.6):
        _, grad = O.scattering(c, np.full_like(c, 0.3), n, 0.0025)
        h = 1e-6
        fd = (O.scattering(c + h, np.full_like(c, 0.3), n, 0.0025)[0] - O.scattering(c - h, np.full_like(c, 0.3), n, 0.0025)[0]) / (2 * h)
        np.testing.assert_allclose(grad, fd, rtol=1e-7)


@pytest.mark.parametrize("pad", [2, 4])
@pytest.mark.parametrize("fresnel_n", [0.0, 1.6])
def test_columns_equal_independent_adc_synthesis(pad, fresnel_n):
    """test_surfel.py:12-20 of the reference, in float64 on both sides."""
    cfg = config(pad)
    positions = aperture_grid(6)
    p = surfels(4)
    amps = np.random.default_rng(0).random(4) + 0.5
    rendered = O.columns(positions, p["centers"], p["normals"], cfg, fresnel_n) @ amps
    synthesized = O.synthesize(positions, p["centers"], p["normals"], amps, cfg, fresnel_n)
    # the only difference is the kernel's +1e-12 m in r: 5e-9 rad of carrier phase
    assert rel_l2(rendered, synthesized) < 2e-8


def test_synthesis_of_an_on_bin_reflector_is_a_single_tone():
    """A reflector exactly on bin 60 (P = 2): D = 1 there and 0 at every other even offset."""
    cfg = config(2)
    chirp = O.Chirp(cfg)
    r = 60 * chirp.spacing
    positions = np.zeros((1, 3), np.float64)
    centers = np.array([[0.0, 0.0, r]])
    normals = np.array([[0.0, 0.0, -1.0]])
    spectrum = O.synthesize(positions, centers, normals, np.ones(1), cfg)[0]
    amp = chirp.lam / (8 * math.pi * r)
    assert abs(abs(spectrum[60]) - amp) < 1e-9 * amp
    even = np.arange(0, 128, 2)
    assert np.abs(spectrum[even[even != 60]]).max() < 1e-9 * amp


def test_render_is_the_columns_times_the_weights():
    cfg = config()
    positions = aperture_grid(4)
    p = surfels(3)
    w = np.array([1 + 0.5j, -0.3j, 2.0])
    cols = O.columns(positions, p["centers"], p["normals"], cfg)
    np.testing.assert_allclose(O.render(positions, p["centers"], p["normals"], w, cfg), cols @ w, atol=0)


@pytest.mark.parametrize("fresnel_n", [0.0, 1.6])
def test_jacobian_matches_central_differences(fresnel_n):
    """test_surfel.py:23-47 of the reference (h = 1e-6)."""
    cfg = config()
    positions = aperture_grid(4).astype(np.float64)
    p = surfels(3)
    centers, normals = p["centers"].astype(np.float64), p["normals"].astype(np.float64)
    _, dcen, dnrm = O.jacobian(positions, centers, normals, cfg, fresnel_n)
    h = 1e-6
    for which, jac in (("c", dcen), ("n", dnrm)):
        for axis in range(3):
            shift = np.zeros_like(centers)
            shift[:, axis] = h
            c1, n1 = (centers + shift, normals) if which == "c" else (centers, normals + shift)
            c2, n2 = (centers - shift, normals) if which == "c" else (centers, normals - shift)
            fd = (O.columns(positions, c1, n1, cfg, fresnel_n) - O.columns(positions, c2, n2, cfg, fresnel_n)) / (2 * h)
            close(jac[..., axis], fd, 2e-5)


def test_jvp_and_vjp_satisfy_finite_differences_and_the_adjoint_identity():
    cfg = config()
    rng = np.random.default_rng(1)
    positions = aperture_grid(5).astype(np.float64)
    p = surfels(4)
    c, n = p["centers"].astype(np.float64), p["normals"].astype(np.float64)
    w = (rng.standard_normal(4) + 1j * rng.standard_normal(4)) * 1e-2
    dc, dn = rng.standard_normal((4, 3)) * 1e-4, rng.standard_normal((4, 3)) * 1e-2
    dw = (rng.standard_normal(4) + 1j * rng.standard_normal(4)) * 1e-2
    jvp = O.jvp(positions, c, n, w, dc, dn, dw, cfg)
    h = 1e-3
    fd = (O.render(positions, c + h * dc, n + h * dn, w + h * dw, cfg) - O.render(positions, c - h * dc, n - h * dn, w - h * dw, cfg)) / (2 * h)
    close(jvp, fd, 1e-4)
    u = rng.standard_normal((25, 128)) + 1j * rng.standard_normal((25, 128))
    gc, gn, gw = O.vjp(positions, c, n, w, u, cfg)
    left = (u.conj() * jvp).real.sum()
    right = (gc * dc).sum() + (gn * dn).sum() + (gw.conj() * dw).real.sum()
    assert abs(left - right) < 1e-10 * abs(left)
    # gradient of L = Re<u, render> along a coordinate by central differences
    for s, j in ((0, 2), (3, 0)):
        e = np.zeros_like(c)
        e[s, j] = 1e-7
        fd_c = ((u.conj() * (O.render(positions, c + e, n, w, cfg) - O.render(positions, c - e, n, w, cfg))).real.sum()) / 2e-7
        assert abs(fd_c - gc[s, j]) < 1e-5 * np.abs(gc).max()


def test_correlations_equal_dictionary_products():
    cfg = config()
    positions = aperture_grid(4)
    p = surfels(5)
    rng = np.random.default_rng(2)
    r = rng.standard_normal((16, 128)) + 1j * rng.standard_normal((16, 128))
    a = O.design(positions, p["centers"], p["normals"], np.ones(5), cfg)
    corr, norm = O.correlate(positions, p["centers"], p["normals"], r, cfg)
    close(corr, a.conj().T @ r.ravel(), 1e-12)
    close(norm, (np.abs(a) ** 2).sum(0), 1e-12)
    real = O.design(positions, p["centers"], p["normals"], np.ones(5), cfg, real=True)
    assert real.shape == (2 * 128 * 16, 5)
    # row (2 k + part) A + a of the real layout
    np.testing.assert_array_equal(real[(2 * 7 + 1) * 16 + 3], a[3 * 128 + 7].imag)


def test_certificate_closed_forms():
    """The certificate is a squared cosine (Cauchy-Schwarz): 1 for a residual along the atom's
    own column, cos^2(phi) in the real mode for that column turned by e^(i phi), 0 for a
    residual orthogonal to it, and in [0, 1] for any residual."""
    cfg = config()
    positions = aperture_grid(4)
    p = surfels(3)
    a = O.columns(positions, p["centers"], p["normals"], cfg)  # [A, K, S]
    rng = np.random.default_rng(5)
    for phi in (0.0, 0.4, 1.2, math.pi / 2):
        r = 2.5 * np.exp(1j * phi) * a[..., 1]
        corr, norm = O.correlate(positions, p["centers"][1:2], p["normals"][1:2], r, cfg)
        np.testing.assert_allclose(O.certificate(corr, norm, r, "complex"), 1.0, rtol=1e-12)
        np.testing.assert_allclose(O.certificate(corr, norm, r, "real"), math.cos(phi) ** 2, rtol=1e-10, atol=1e-15)
    column = a[..., 0].ravel()
    noise = rng.standard_normal(column.shape) + 1j * rng.standard_normal(column.shape)
    orthogonal = (noise - column * (column.conj() @ noise) / (column.conj() @ column)).reshape(a.shape[:2])
    corr, norm = O.correlate(positions, p["centers"][:1], p["normals"][:1], orthogonal, cfg)
    assert O.certificate(corr, norm, orthogonal, "complex")[0] < 1e-24
    r = noise.reshape(a.shape[:2])
    corr, norm = O.correlate(positions, p["centers"], p["normals"], r, cfg)
    value = O.certificate(corr, norm, r, "complex")
    assert np.all((value >= 0) & (value <= 1))


def test_spectral_atoms_equal_a_brute_force_windowed_dft():
    """A 2-D atom is the product of per-axis tapped DFT sums (pins oracle.spectral)."""
    from oracle import spectral as Sp

    shape, samples = (24, 20), (16, 20)
    centers = np.array([[3.3, 4.1], [12.0, 15.0], [18.4, 7.9]])
    bins = Sp.grid_bins(shape)
    for window in ("rect", "hann", "hamming", "kaiser"):
        a, da = Sp.design(shape, samples, window, 6.0, 0.0, centers, bins)
        expected = np.ones_like(a)
        for d in range(2):
            delta = bins[:, None, d] - centers[None, :, d]
            expected = expected * K.dirichlet_bruteforce(delta, samples[d], shape[d], K.window_taps(window, samples[d], 6.0))[0]
        assert np.abs(a - expected).max() < 1e-12
        h = 1e-6
        shifted = centers.copy()
        shifted[:, 0] += h
        fd = (Sp.design(shape, samples, window, 6.0, 0.0, shifted, bins)[0] - a) / h
        assert np.abs(da[..., 0] - fd).max() < 1e-4 * np.abs(fd).max()


def test_search_oracles_on_known_answers():
    from oracle import search as Se

    centers = np.array([[0.0, 0, 0], [0.05, 0, 0], [1.0, 0, 0], [1.02, 0, 0]])
    scores = np.array([0.5, 0.9, 0.9, 0.1])
    # index 1 wins the tie with 2 by index; it suppresses 0; then 2; 3 is suppressed by 2
    assert Se.peaks(scores, centers, 4, 0.1) == [1, 2]
    assert Se.peaks(scores, centers, 4, 0.0) == [1, 2, 0, 3]
    pos = np.zeros((1, 3), np.float32)
    pts = np.array([[0, 0, 2.5], [0, 0, 3.5], [0, 0, 4.2]], np.float32)  # bins 2.5 -> 2, 3.5 -> 4, 4.2 -> 4
    intervals = np.zeros((1, 8), np.int64)
    intervals[0, 4] = 1
    assert Se.range_gate(pts, pos, intervals, 1.0, 1).tolist() == [False, True, True]


def test_range_response_and_energy():
    from oracle import linalg as L

    r = L.range_response(60.0, 70.0, 11, 128, 128, 256)
    assert r.shape == (128, 11)
    assert abs(r[60, 0] - 1) < 1e-15 and abs(r[70, 10] - 1) < 1e-15
    assert L.energy([1 + 1j, 2], [1, 0]) == 5.0


def _spd(rng, count, width):
    """count random symmetric positive-definite [width, width] matrices."""
    m = rng.standard_normal((count, width, width))
    return np.einsum("cij,ckj->cik", m, m) + width * np.eye(width)[None]


def test_search_direction_is_the_newton_step_of_a_quadratic():
    """With inverse = H^-1 of a quadratic f(x) = 1/2 x^T H x - b^T x, one unclipped step from
    x lands on the minimiser H^-1 b (solved by LAPACK, not by the oracle's formula)."""
    from oracle import search as Se

    rng = np.random.default_rng(11)
    h = _spd(rng, 5, 3)
    b = rng.standard_normal((5, 3))
    x = rng.standard_normal((5, 3))
    g = np.einsum("cij,cj->ci", h, x) - b
    d = Se.direction(np.linalg.inv(h), g, np.zeros((5, 3)), np.ones(3), 1e9)
    np.testing.assert_allclose(x + d, np.linalg.solve(h, b[..., None])[..., 0], rtol=1e-10, atol=1e-12)


def test_search_direction_fallbacks_mask_tangent_and_radius():
    from oracle import search as Se

    rng = np.random.default_rng(12)
    g = rng.standard_normal((6, 3))
    # identity inverse: steepest descent, clipped to the trust radius (length exactly radius)
    d = Se.direction(np.broadcast_to(np.eye(3), (6, 3, 3)), g, np.zeros((6, 3)), np.ones(3), 0.25)
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 0.25, rtol=1e-12)
    np.testing.assert_allclose(d / np.linalg.norm(d, axis=1, keepdims=True), -g / np.linalg.norm(g, axis=1, keepdims=True), rtol=1e-12)
    # a negated (ascent) inverse: d . g >= 0, so the step falls back to -g
    d = Se.direction(np.broadcast_to(-np.eye(3), (6, 3, 3)), g, np.zeros((6, 3)), np.ones(3), 1e9)
    np.testing.assert_allclose(d, -g, rtol=1e-12)
    # masked components stay exactly zero
    mask = np.array([1.0, 0.0, 1.0])
    d = Se.direction(_spd(rng, 6, 3), g, np.zeros((6, 3)), mask, 1e9)
    assert np.all(d[:, 1] == 0)
    # width 6: the normal part is tangent to the unit normal
    n = rng.standard_normal((6, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    g6 = rng.standard_normal((6, 6))
    d = Se.direction(_spd(rng, 6, 6), g6, n, np.ones(6), 1e9)
    np.testing.assert_allclose((d[:, 3:] * n).sum(1), 0, atol=1e-12)
    # a length below the radius is not scaled
    small = Se.direction(np.broadcast_to(np.eye(3), (6, 3, 3)), 1e-3 * g, np.zeros((6, 3)), np.ones(3), 1.0)
    np.testing.assert_allclose(small, -1e-3 * g, rtol=1e-12)


def test_search_trial_clamps_renormalises_and_freezes():
    from oracle import search as Se

    rng = np.random.default_rng(13)
    c = rng.uniform(-1, 1, (7, 3))
    n = rng.standard_normal((7, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    d = rng.standard_normal((7, 6))
    controls = np.array([
        [2.0, 0.5, 1.0, 1.0, 1.0, 1.0],     # per-coordinate scale
        [1.0, 1.0, 0.0, 1.0, 1.0, 1.0],     # trainable: the z centre coordinate is frozen
        [-0.5, -0.5, -0.5, 0, 0, 0],        # lower bounds
        [0.5, 0.5, 0.5, 0, 0, 0],           # upper bounds
    ])  # fmt: skip
    oc, on = Se.trial(c, n, d, controls, 0.5)
    assert np.all(oc[:, 2] == c[:, 2])
    assert np.all(oc[:, :2] >= -0.5) and np.all(oc[:, :2] <= 0.5)
    # an interior move is exactly centre + fraction * scale * direction
    inside = np.all(np.abs(c[:, :2] + 0.5 * controls[0, :2] * d[:, :2]) < 0.5, axis=1)
    np.testing.assert_allclose(oc[inside, :2], (c[:, :2] + 0.5 * controls[0, :2] * d[:, :2])[inside], rtol=1e-12)
    np.testing.assert_allclose(np.linalg.norm(on, axis=1), 1, rtol=1e-12)
    # the new normal points along n + fraction * d_n
    m = n + 0.5 * d[:, 3:]
    np.testing.assert_allclose(np.cross(on, m), 0, atol=1e-12)
    assert np.all((on * m).sum(1) > 0)
    # width 3: normals untouched
    _, same = Se.trial(c, n, d[:, :3], controls, 0.5)
    assert np.array_equal(same, n)
    # a zero fraction returns the (clamped) start
    oc0, on0 = Se.trial(c, n, d, controls, 0.0)
    np.testing.assert_allclose(oc0, np.where(controls[1, :3] > 0, np.clip(c, -0.5, 0.5), c))
    np.testing.assert_allclose(on0, n, rtol=1e-12)
