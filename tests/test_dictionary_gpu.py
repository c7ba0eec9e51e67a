"""Dictionary atoms, the explicit Jacobian and the correlations on the GPU vs the float64 oracle."""

import numpy as np
import pytest
import torch
from conftest import aperture_grid, rel_l2, surfels

from oracle import surfel as O

pytestmark = pytest.mark.gpu
TOL = 1e-5


def f32(x):
    return torch.as_tensor(np.asarray(x, np.float32)).cuda()


def scene(count=11, seed=3, side=6):
    p = surfels(count, seed=seed)
    return aperture_grid(side), p["centers"], p["normals"], p["areas"]


@pytest.mark.parametrize("pad,fresnel_n", [(2, 0.0), (4, 1.6), (1, 0.0), (8, 0.0)])
def test_design_matches_the_oracle_in_both_layouts(pad, fresnel_n):
    from dsplat import FMCWConfig
    from dsplat.kernels import surfel

    cfg = FMCWConfig(pad_factor=pad)
    positions, centers, normals, areas = scene()
    areas = areas * np.linspace(0.5, 2.0, len(areas))
    for real in (False, True):
        got = surfel.design(cfg, f32(positions), f32(centers), f32(normals), f32(areas), fresnel_n, real)
        want = O.design(positions, centers, normals, areas, cfg, fresnel_n, real)
        assert got.shape == want.shape
        assert rel_l2(got.cpu().numpy(), want) < TOL


@pytest.mark.parametrize("fresnel_n", [0.0, 1.6])
def test_jacobian_matches_the_oracle(fresnel_n):
    from dsplat import FMCWConfig
    from dsplat.kernels import surfel

    cfg = FMCWConfig()
    positions, centers, normals, _ = scene(7, 1, 4)
    col, dc, dn = surfel.jacobian(cfg, f32(positions), f32(centers), f32(normals), fresnel_n)
    wcol, wdc, wdn = O.jacobian(positions, centers, normals, cfg, fresnel_n)
    assert rel_l2(col.cpu().numpy(), wcol) < TOL
    assert rel_l2(dc.cpu().numpy(), wdc) < TOL
    assert rel_l2(dn.cpu().numpy(), wdn) < TOL


@pytest.mark.parametrize("pad,fresnel_n", [(2, 0.0), (2, 1.6), (4, 0.0), (1, 0.0)])
def test_correlations_match_the_oracle(pad, fresnel_n):
    from dsplat import FMCWConfig
    from dsplat.kernels import surfel

    cfg = FMCWConfig(pad_factor=pad)
    positions, centers, normals, _ = scene(37, 5, 8)
    rng = np.random.default_rng(pad)
    res = rng.standard_normal((64, cfg.num_range_bins)) + 1j * rng.standard_normal((64, cfg.num_range_bins))
    r = torch.as_tensor(res.astype(np.complex64)).cuda()
    corr, norm = surfel.correlate(cfg, f32(positions), f32(centers), f32(normals), r, fresnel_n)
    wc, wn = O.correlate(positions, centers, normals, res, cfg, fresnel_n)
    assert rel_l2(corr.cpu().numpy(), wc) < TOL and rel_l2(norm.cpu().numpy(), wn) < TOL
    for include_normals in (False, True):
        c, n, dc, dn = surfel.correlate_gradient(cfg, f32(positions), f32(centers), f32(normals), r, fresnel_n,
                                                 include_normals)
        ec, en, edc, edn = O.correlate_gradient(positions, centers, normals, res, cfg, fresnel_n, include_normals)
        for got, want in ((c, ec), (n, en), (dc, edc), (dn, edn)):
            assert rel_l2(got.cpu().numpy(), want) < TOL
    bank = np.asarray(torch.nn.functional.normalize(torch.randn(6, 3, generator=torch.Generator().manual_seed(1)), dim=1))
    oc, on = surfel.correlate_orientations(cfg, f32(positions), f32(centers), f32(normals), f32(bank), r, fresnel_n)
    ec, en = O.correlate_orientations(positions, centers, normals, bank, res, cfg, fresnel_n)
    assert oc.shape == (37, 7)
    assert rel_l2(oc.cpu().numpy(), ec) < TOL and rel_l2(on.cpu().numpy(), en) < TOL


def test_correlate_is_the_adjoint_of_the_design():
    """a^H r and |a|^2 from the GPU correlate equal products with the GPU design."""
    from dsplat import FMCWConfig
    from dsplat.kernels import surfel

    cfg = FMCWConfig()
    positions, centers, normals, areas = scene(19, 8, 8)
    r = torch.randn(64, 128, dtype=torch.complex64, generator=torch.Generator().manual_seed(0)).cuda()
    a = surfel.design(cfg, f32(positions), f32(centers), f32(normals), torch.ones(19, device="cuda"), 0.0, False)
    corr, norm = surfel.correlate(cfg, f32(positions), f32(centers), f32(normals), r, 0.0)
    a64 = a.to(torch.complex128)
    assert rel_l2(corr.cpu().numpy(), (a64.conj().T @ r.flatten().to(torch.complex128)).cpu().numpy()) < 1e-6
    assert rel_l2(norm.cpu().numpy(), a64.abs().square().sum(0).cpu().numpy()) < 1e-6


@pytest.mark.parametrize("samples,fresnel_n", [(128, 0.0), (256, 1.6)])
def test_adjoint_correlation_matches_the_oracle(samples, fresnel_n):
    """A^H r through the sample-domain VJP (the batched right-hand sides), with area, against
    the float64 oracle's dense columns, at an atom count that takes that path."""
    from dsplat import FMCWConfig, FMCWModel, Scan

    cfg = FMCWConfig(adc_samples=samples)
    positions, centers, normals, areas = scene(count=160, side=6)
    areas = areas * np.linspace(0.5, 2.0, 160)
    rng = np.random.default_rng(samples)
    r = rng.standard_normal((36, cfg.num_range_bins)) + 1j * rng.standard_normal((36, cfg.num_range_bins))
    model = FMCWModel(Scan(torch.as_tensor(r, dtype=torch.complex64).cuda(), f32(positions), cfg), fresnel_n=fresnel_n)
    params = dict(centers=f32(centers), normals=f32(normals), areas=f32(areas))
    assert len(centers) >= model.ADJOINT_MIN_ATOMS
    got = model.adjoint_correlation(params, model.scan.data).cpu().numpy()
    a = O.columns(positions, centers, normals, cfg, fresnel_n) * np.sqrt(areas)
    want = a.reshape(-1, len(centers)).conj().T @ r.reshape(-1)
    assert rel_l2(got, want) < TOL


@pytest.mark.parametrize("samples,pad,fresnel_n", [(128, 2, 0.0), (128, 4, 1.6), (256, 2, 0.0), (256, 2, 1.6)])
def test_sample_domain_certificate_gradient_matches_the_oracle(samples, pad, fresnel_n):
    """Correlations, norms and their centre and normal derivatives from the sample-domain
    certificate operands (two candidate tiles, ragged) against the float64 oracle."""
    from dsplat import FMCWConfig
    from dsplat.kernels import surfel

    cfg = FMCWConfig(adc_samples=samples, pad_factor=pad)
    positions, centers, normals, _ = scene(count=150, side=6)
    rng = np.random.default_rng(samples + pad)
    r = rng.standard_normal((36, cfg.num_range_bins)) + 1j * rng.standard_normal((36, cfg.num_range_bins))
    ops = surfel.CertificateOperands(cfg, torch.as_tensor(r, dtype=torch.complex64).cuda(), fresnel_n)
    out = surfel.certificate_gradient(f32(positions), f32(centers), f32(normals), ops).double().cpu().numpy()
    ec, en, edc, edn = O.correlate_gradient(positions, centers, normals, r, cfg, fresnel_n, True)
    assert rel_l2(out[:, 0] + 1j * out[:, 1], ec) < TOL
    assert rel_l2(out[:, 2], en) < TOL
    assert rel_l2(out[:, 3:9] + 1j * out[:, 9:15], edc) < TOL
    assert rel_l2(out[:, 15:21], edn) < TOL
