"""Surfel render and VJP on the GPU against the float64 oracle.

Bar: 1e-5 relative L2 against float64 for forward and VJP. The reference's own float32
kernels sit near 2e-4 (carrier phase), so this is the stricter bar.
"""

import sys

import numpy as np
import pytest
import torch
from conftest import aperture_grid, rel_l2, surfels

from oracle import surfel as O

pytestmark = pytest.mark.gpu
TOL = 1e-5


def dev(x, dtype=None):
    t = torch.as_tensor(np.asarray(x))
    return t.to("cuda", dtype) if dtype is not None else t.cuda()


def say(*args):
    print(*args, file=sys.stderr, flush=True)


def run_render(cfg, positions, centers, normals, weights, fresnel_n=0.0):
    from dsplat.kernels import surfel

    out = surfel.render(
        cfg, dev(positions, torch.float32), dev(centers, torch.float32), dev(normals, torch.float32),
        dev(weights, torch.complex64), fresnel_n,
    )  # fmt: skip
    torch.cuda.synchronize()
    return out.cpu().numpy()


def run_vjp(cfg, positions, centers, normals, weights, upstream, fresnel_n=0.0):
    from dsplat.kernels import surfel

    gc, gn, gw = surfel.vjp(
        cfg, dev(positions, torch.float32), dev(centers, torch.float32), dev(normals, torch.float32),
        dev(weights, torch.complex64), dev(upstream, torch.complex64), fresnel_n,
    )  # fmt: skip
    torch.cuda.synchronize()
    return gc.cpu().numpy(), gn.cpu().numpy(), gw.cpu().numpy()


def f32(x):
    return np.asarray(x, np.float32)


def test_render_and_vjp_match_the_oracle_on_the_small_fmcw_case():
    from dsplat import workloads

    w = workloads.fmcw_small()
    args = (w["config"], w["positions"].numpy(), w["centers"].numpy(), w["normals"].numpy())
    weights = w["weights"].numpy()
    say("render 1024x128x200")
    out = run_render(*args, weights)
    expected = O.render(*args[1:], weights, args[0])
    err = rel_l2(out, expected)
    say(f"render rel L2 {err:.3e}")
    assert err < TOL
    # VJP on the first 256 apertures (the oracle's Jacobian is the cost here)
    sub = (args[0], args[1][:256], args[2], args[3])
    up = w["cotangent"].numpy()[:256]
    say("vjp 256x128x200")
    gc, gn, gw = run_vjp(*sub, weights, up)
    ec, en, ew = O.vjp(*sub[1:], weights, up, sub[0])
    errs = rel_l2(gc, ec), rel_l2(gn, en), rel_l2(gw, ew)
    say("vjp rel L2 centers %.3e normals %.3e weights %.3e" % errs)
    assert max(errs) < TOL


@pytest.mark.parametrize("pad,fresnel_n", [(2, 1.6), (4, 0.0), (4, 1.6), (1, 0.0), (8, 0.0)])
def test_render_and_vjp_across_pad_factors_and_materials(pad, fresnel_n):
    from dsplat import FMCWConfig

    cfg = FMCWConfig(pad_factor=pad)
    positions = aperture_grid(8)
    p = surfels(12, seed=3)
    rng = np.random.default_rng(pad)
    weights = (rng.standard_normal(12) + 1j * rng.standard_normal(12)) * 1e-2
    say(f"render pad {pad} n {fresnel_n}")
    out = run_render(cfg, positions, p["centers"], p["normals"], weights, fresnel_n)
    expected = O.render(positions, p["centers"], p["normals"], weights, cfg, fresnel_n)
    assert rel_l2(out, expected) < TOL
    up = rng.standard_normal((64, cfg.num_range_bins)) + 1j * rng.standard_normal((64, cfg.num_range_bins))
    got = run_vjp(cfg, positions, p["centers"], p["normals"], weights, up, fresnel_n)
    want = O.vjp(positions, p["centers"], p["normals"], weights, up, cfg, fresnel_n)
    for g, e in zip(got, want, strict=True):
        assert rel_l2(g, e) < TOL


def test_lobe_edge_cases():
    """A reflector exactly on a bin, one whose peak wraps past the observed half, and a
    back face. The kernels treat the lobe bin apart; these are its corner cases."""
    from dsplat import FMCWConfig

    cfg = FMCWConfig()
    spacing = cfg.range_bin_spacing_m
    positions = np.zeros((2, 3), np.float32)
    positions[1, 0] = 0.003
    ranges = [60 * spacing, 60.5 * spacing, 200.25 * spacing, 255.9 * spacing, 0.3]
    centers = f32([[0, 0, r] for r in ranges])
    normals = f32([[0, 0, -1]] * 4 + [[0, 0, 1]])  # the last one faces away
    weights = np.array([1.0, 0.7j, 0.5, -0.4, 1.0])
    say("lobe edge cases")
    out = run_render(cfg, positions, centers, normals, weights)
    expected = O.render(positions, centers, normals, weights, cfg)
    assert rel_l2(out, expected) < TOL
    up = np.random.default_rng(5).standard_normal((2, 128)) + 0j
    got = run_vjp(cfg, positions, centers, normals, weights, up)
    want = O.vjp(positions, centers, normals, weights, up, cfg)
    for g, e in zip(got, want, strict=True):
        assert rel_l2(g, e) < TOL
    assert np.all(got[0][4] == 0) and got[2][4] == 0  # the back face gets no gradient


@pytest.mark.parametrize("samples,fresnel_n", [(128, 0.0), (256, 1.6)])
def test_vjp_with_banded_and_wide_apertures(samples, fresnel_n):
    """The sample-domain VJP forms each aperture's pair sums from moments of a band of at most
    32 bins around the range interval to the centres' box; an aperture that sees the box over
    more bins takes the exact sum over the samples. Apertures along a 1 m line see the same
    surfels over 4 to ~40 bins, so one call mixes both."""
    from dsplat import FMCWConfig

    cfg = FMCWConfig(adc_samples=samples)
    positions = np.zeros((16, 3), np.float32)
    positions[:, 0] = np.linspace(0.0, 1.0, 16, dtype=np.float32)
    rng = np.random.default_rng(11)
    centers = f32(np.stack([rng.uniform(-0.1, 0.1, 24), rng.uniform(-0.02, 0.02, 24), rng.uniform(0.3, 0.31, 24)], -1))
    normals = f32(np.tile([0.0, 0.0, -1.0], (24, 1)) + 0.3 * rng.standard_normal((24, 3)))
    normals /= np.linalg.norm(normals, axis=1, keepdims=True)
    weights = (rng.standard_normal(24) + 1j * rng.standard_normal(24)) * 1e-2
    up = rng.standard_normal((16, cfg.num_range_bins)) + 1j * rng.standard_normal((16, cfg.num_range_bins))
    say(f"vjp banded + wide apertures, {samples} samples, n {fresnel_n}")
    got = run_vjp(cfg, positions, centers, normals, weights, up, fresnel_n)
    want = O.vjp(positions, centers, normals, weights, up, cfg, fresnel_n)
    errs = [rel_l2(g, e) for g, e in zip(got, want, strict=True)]
    say("  rel L2 centers %.3e normals %.3e weights %.3e" % tuple(errs))
    assert max(errs) < TOL


def test_autograd_backward_is_the_vjp_and_is_deterministic():
    from dsplat.kernels import surfel
    from dsplat import FMCWConfig

    cfg = FMCWConfig()
    positions = dev(aperture_grid(8), torch.float32)
    p = surfels(9, seed=4)
    centers = dev(p["centers"], torch.float32).requires_grad_()
    normals = dev(p["normals"], torch.float32).requires_grad_()
    weights = (torch.linspace(0.5, 1.5, 9, device="cuda") * 1e-2).to(torch.complex64).requires_grad_()
    up = torch.randn(64, 128, dtype=torch.complex64, device="cuda", generator=torch.Generator("cuda").manual_seed(0))
    (up.conj() * surfel.render(cfg, positions, centers, normals, weights, 0.0)).real.sum().backward()
    gc, gn, gw = surfel.vjp(cfg, positions, centers.detach(), normals.detach(), weights.detach(), up, 0.0)
    assert torch.equal(centers.grad, gc) and torch.equal(normals.grad, gn) and torch.equal(weights.grad, gw)
    again = surfel.vjp(cfg, positions, centers.detach(), normals.detach(), weights.detach(), up, 0.0)
    assert all(torch.equal(a, b) for a, b in zip(again, (gc, gn, gw), strict=True))


def test_empty_inputs():
    from dsplat.kernels import surfel
    from dsplat import FMCWConfig

    cfg = FMCWConfig()
    positions = dev(aperture_grid(2), torch.float32)
    none = torch.zeros(0, 3, device="cuda")
    out = surfel.render(cfg, positions, none, none, torch.zeros(0, dtype=torch.complex64, device="cuda"), 0.0)
    assert out.shape == (4, 128) and torch.count_nonzero(out) == 0
    p = surfels(3)
    c, n = dev(p["centers"], torch.float32), dev(p["normals"], torch.float32)
    w = torch.ones(3, dtype=torch.complex64, device="cuda")
    assert surfel.render(cfg, torch.zeros(0, 3, device="cuda"), c, n, w, 0.0).shape == (0, 128)
    gc, gn, gw = surfel.vjp(cfg, torch.zeros(0, 3, device="cuda"), c, n, w, torch.zeros(0, 128, dtype=torch.complex64, device="cuda"), 0.0)
    assert torch.count_nonzero(gc) == 0 and torch.count_nonzero(gw) == 0


def test_full_size_render_and_vjp():
    """128^2 apertures x 2048 ellipsoid surfels: aperture-subset parity plus the
    size-independent adjoint identity Re<u, A w> = Re<A^H u, w>."""
    from dsplat import workloads
    from dsplat.kernels import surfel

    w = workloads.ellipsoid_scene()
    cfg = w["config"]
    positions, centers, normals = (w[k].cuda() for k in ("positions", "centers", "normals"))
    weights, up = w["weights"].cuda(), w["cotangent"].cuda()
    say("full render 16384x128x2048")
    out = surfel.render(cfg, positions, centers, normals, weights, 0.0)
    gc, gn, gw = surfel.vjp(cfg, positions, centers, normals, weights, up, 0.0)
    torch.cuda.synchronize()
    left = float((up.to(torch.complex128).conj() * out.to(torch.complex128)).real.sum())
    right = float((gw.to(torch.complex128).conj() * weights.to(torch.complex128)).real.sum())
    assert abs(left - right) < 1e-5 * abs(left)
    # every 256th aperture against the oracle
    ids = np.arange(0, 16384, 256)
    pos_np, c_np, n_np, w_np = w["positions"].numpy(), w["centers"].numpy(), w["normals"].numpy(), w["weights"].numpy()
    expected = O.render(pos_np[ids], c_np, n_np, w_np, cfg)
    err = rel_l2(out.cpu().numpy()[ids], expected)
    say(f"full render subset rel L2 {err:.3e}")
    assert err < TOL
    sub = surfel.vjp(cfg, positions[ids], centers, normals, weights, up[ids], 0.0)
    want = O.vjp(pos_np[ids], c_np, n_np, w_np, w["cotangent"].numpy()[ids], cfg)
    errs = [rel_l2(g.cpu().numpy(), e) for g, e in zip(sub, want, strict=True)]
    say("full vjp subset rel L2 %s" % ["%.3e" % e for e in errs])
    assert max(errs) < TOL


def run_jvp(cfg, positions, centers, normals, weights, dc, dn, dw, fresnel_n=0.0):
    from dsplat.kernels import surfel

    out = surfel.jvp(
        cfg, dev(positions, torch.float32), dev(centers, torch.float32), dev(normals, torch.float32),
        dev(weights, torch.complex64), dev(dc, torch.float32), dev(dn, torch.float32),
        dev(dw, torch.complex64), fresnel_n,
    )  # fmt: skip
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("pad,fresnel_n", [(2, 0.0), (2, 1.6), (4, 0.0), (1, 0.0), (8, 1.6)])
def test_jvp_matches_the_oracle(pad, fresnel_n):
    from dsplat import FMCWConfig

    cfg = FMCWConfig(pad_factor=pad)
    positions = aperture_grid(8)
    p = surfels(10, seed=6)
    rng = np.random.default_rng(10 + pad)
    w = (rng.standard_normal(10) + 1j * rng.standard_normal(10)) * 1e-2
    dc, dn = f32(rng.standard_normal((10, 3)) * 1e-4), f32(rng.standard_normal((10, 3)) * 1e-2)
    dw = (rng.standard_normal(10) + 1j * rng.standard_normal(10)) * 1e-2
    say(f"jvp pad {pad} n {fresnel_n}")
    got = run_jvp(cfg, positions, p["centers"], p["normals"], w, dc, dn, dw, fresnel_n)
    want = O.jvp(positions, p["centers"], p["normals"], w, dc, dn, dw, cfg, fresnel_n)
    err = rel_l2(got, want)
    say(f"jvp rel L2 {err:.3e}")
    assert err < TOL


def test_jvp_lobe_edge_cases_and_the_adjoint_identity():
    from dsplat import FMCWConfig

    cfg = FMCWConfig()
    spacing = cfg.range_bin_spacing_m
    positions = np.zeros((3, 3), np.float32)
    positions[1, 0], positions[2, 1] = 0.003, -0.002
    ranges = [60 * spacing, 60.5 * spacing, 200.25 * spacing, 255.9 * spacing, 0.3, 70.0001 * spacing]
    centers = f32([[0, 0, r] for r in ranges])
    normals = f32([[0, 0, -1]] * 4 + [[0, 0, 1], [0.1, 0, -1]])
    normals = f32(normals / np.linalg.norm(normals, axis=1, keepdims=True))
    rng = np.random.default_rng(7)
    w = rng.standard_normal(6) + 1j * rng.standard_normal(6)
    dc, dn = f32(rng.standard_normal((6, 3)) * 1e-4), f32(rng.standard_normal((6, 3)) * 1e-2)
    dw = rng.standard_normal(6) + 1j * rng.standard_normal(6)
    got = run_jvp(cfg, positions, centers, normals, w, dc, dn, dw)
    want = O.jvp(positions, centers, normals, w, dc, dn, dw, cfg)
    assert rel_l2(got, want) < TOL
    up = rng.standard_normal((3, 128)) + 1j * rng.standard_normal((3, 128))
    gc, gn, gw = run_vjp(cfg, positions, centers, normals, w, up)
    left = (up.conj() * got).real.sum()
    right = (gc * dc).sum() + (gn * dn).sum() + (gw.conj() * dw).real.sum()
    assert abs(left - right) <= 1e-5 * abs(left)


def test_forward_mode_autograd_uses_the_jvp_kernel():
    import torch.autograd.forward_ad as fwAD

    from dsplat import FMCWConfig
    from dsplat.kernels import surfel

    cfg = FMCWConfig()
    positions = dev(aperture_grid(8), torch.float32)
    p = surfels(7, seed=8)
    c, n = dev(p["centers"], torch.float32), dev(p["normals"], torch.float32)
    w = torch.full((7,), 0.01 + 0.005j, dtype=torch.complex64, device="cuda")
    g = torch.Generator("cuda").manual_seed(3)
    dc = torch.randn(7, 3, device="cuda", generator=g) * 1e-4
    dn = torch.randn(7, 3, device="cuda", generator=g) * 1e-2
    dw = torch.randn(7, dtype=torch.complex64, device="cuda", generator=g) * 1e-3
    with fwAD.dual_level():
        out = surfel.render(cfg, positions, fwAD.make_dual(c, dc), fwAD.make_dual(n, dn), fwAD.make_dual(w, dw), 0.0)
        tangent = fwAD.unpack_dual(out).tangent
    direct = surfel.jvp(cfg, positions, c, n, w, dc, dn, dw, 0.0)
    torch.testing.assert_close(tangent, direct, rtol=0, atol=0)


def test_full_size_jvp_against_finite_differences_of_the_render():
    """The 128^2 x 2048 scene: JVP against central differences of the HIP render with
    h = 1e-6 m on centers and 1e-4 on normals, and the aperture-subset oracle."""
    from dsplat import workloads
    from dsplat.kernels import surfel

    w = workloads.ellipsoid_scene()
    cfg = w["config"]
    positions, centers, normals, weights = (w[k].cuda() for k in ("positions", "centers", "normals", "weights"))
    g = torch.Generator().manual_seed(2)
    dc = torch.randn(len(centers), 3, generator=g).cuda()
    dn = torch.randn(len(centers), 3, generator=g).cuda()
    zero_c, zero_w = torch.zeros_like(dc), torch.zeros_like(weights)
    say("full jvp vs FD")
    # float32 render noise (~3e-7 |y|) over the FD signal: 4e-5 for centers (phase
    # moves 5e-3 rad per step), 1e-3 for normals (amplitude moves 2e-4)
    for delta_c, delta_n, h, tol in ((dc, zero_c, 1e-6, 1e-3), (zero_c, dn, 1e-4, 1e-2)):
        cp, cm = centers + h * delta_c, centers - h * delta_c
        np_, nm = normals + h * delta_n, normals - h * delta_n
        # float32 coordinates quantize the step (ulp(0.35 m) = 3e-8 m): differentiate
        # along the step that was actually taken
        eff_c = ((cp.double() - cm.double()) / (2 * h)).float()
        eff_n = ((np_.double() - nm.double()) / (2 * h)).float()
        jvp = surfel.jvp(cfg, positions, centers, normals, weights, eff_c, eff_n, zero_w, 0.0)
        plus = surfel.render(cfg, positions, cp, np_, weights, 0.0)
        minus = surfel.render(cfg, positions, cm, nm, weights, 0.0)
        fd = (plus.to(torch.complex128) - minus.to(torch.complex128)) / (2 * h)
        err = float((jvp.to(torch.complex128) - fd).norm() / fd.norm())
        say(f"h {h}: jvp vs FD rel L2 {err:.3e}")
        assert err < tol
    ids = np.arange(0, 16384, 512)
    # float64 central differences of the oracle at h = 1e-6 m (centers), 1e-4 (normals)
    pos_np, c_np, n_np, w_np = (w[k].numpy() for k in ("positions", "centers", "normals", "weights"))
    c64, n64 = c_np.astype(np.float64), n_np.astype(np.float64)
    for which, h in (("c", 1e-6), ("n", 1e-4)):
        d = dc.cpu().numpy().astype(np.float64) if which == "c" else dn.cpu().numpy().astype(np.float64)
        dcc, dnn = (d, np.zeros_like(d)) if which == "c" else (np.zeros_like(d), d)
        plus = O.render(pos_np[ids], c64 + h * dcc, n64 + h * dnn, w_np, cfg)
        minus = O.render(pos_np[ids], c64 - h * dcc, n64 - h * dnn, w_np, cfg)
        got = surfel.jvp(cfg, positions[ids], centers, normals, weights, torch.as_tensor(dcc, dtype=torch.float32).cuda(),
                         torch.as_tensor(dnn, dtype=torch.float32).cuda(), zero_w, 0.0)
        err = rel_l2(got.cpu().numpy(), (plus - minus) / (2 * h))
        say(f"h {h}: jvp vs float64 FD of the oracle rel L2 {err:.3e}")
        assert err < 1e-4
    dcs, dns, dws = (w[k] for k in ("dcenters", "dnormals", "dweights"))
    got = surfel.jvp(cfg, positions[ids], centers, normals, weights, dcs.cuda(), dns.cuda(), dws.cuda(), 0.0)
    want = O.jvp(w["positions"].numpy()[ids], w["centers"].numpy(), w["normals"].numpy(), w["weights"].numpy(),
                 dcs.numpy(), dns.numpy(), dws.numpy(), cfg)
    err = rel_l2(got.cpu().numpy(), want)
    say(f"full jvp subset rel L2 {err:.3e}")
    assert err < TOL


@pytest.mark.parametrize("samples", [128, 256])
def test_full_size_vjp_blocks_by_the_adjoint_identity_and_aperture_additivity(samples):
    """128^2 apertures x 2048 surfels: every gradient block of the full-size VJP (centres,
    normals, weights, each with its own tangent) against the JVP by the adjoint identity
    Re<u, J v> = <gc, dc> + <gn, dn> + Re<gw, dw> (to 1e-5 of the larger of the inner product and
    1e-2 of the sum of its terms' magnitudes), and the full-size VJP against the sum of the
    VJPs of four aperture quarters (the gradient is additive over apertures; each quarter
    reduces over its own aperture ranges)."""
    from dsplat import FMCWConfig, workloads
    from dsplat.kernels import surfel

    w = workloads.ellipsoid_scene()
    cfg = FMCWConfig(adc_samples=samples)
    K = cfg.num_range_bins
    positions, centers, normals, weights = (w[k].cuda() for k in ("positions", "centers", "normals", "weights"))
    up = workloads.complex_normal((len(positions), K), torch.Generator().manual_seed(1)).cuda()
    say(f"full-size vjp blocks, {samples} samples")
    gc, gn, gw = surfel.vjp(cfg, positions, centers, normals, weights, up, 0.0)
    dcs, dns, dws = (w[k].cuda() for k in ("dcenters", "dnormals", "dweights"))
    zc, zw = torch.zeros_like(dcs), torch.zeros_like(dws)
    u64 = up.to(torch.complex128)
    for name, (dc, dn, dw) in {"centers": (dcs, zc, zw), "normals": (zc, dns, zw), "weights": (zc, zc, dws)}.items():
        jv = surfel.jvp(cfg, positions, centers, normals, weights, dc, dn, dw, 0.0).to(torch.complex128)
        terms = (u64.conj() * jv).real
        # the inner product's own scale: the sum of its terms' magnitudes (the normals' sum
        # cancels to ~1e-3 of it at 256 samples)
        left, scale = float(terms.sum()), float(terms.abs().sum())
        right = float((gc.double() * dc.double()).sum() + (gn.double() * dn.double()).sum()
                      + (gw.to(torch.complex128).conj() * dw.to(torch.complex128)).real.sum())  # fmt: skip
        say(f"  {name}: <u, J v> {left:.6e} vs <g, v> {right:.6e} (scale {scale:.3e})")
        assert abs(left - right) < 1e-5 * max(abs(left), 1e-2 * scale)
    parts = [surfel.vjp(cfg, positions[q], centers, normals, weights, up[q], 0.0) for q in torch.arange(16384).cuda().chunk(4)]
    for full, pieces in zip((gc, gn, gw), zip(*parts)):
        total = sum(p.to(torch.complex128 if p.is_complex() else torch.float64) for p in pieces)
        err = float((full.to(total.dtype) - total).norm() / total.norm())
        assert err < 1e-6, err
