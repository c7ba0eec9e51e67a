"""The batched certificate scan (several scans of uneven size in one launch) on the GPU
against the float64 oracle's correlation, the single-scan correlate kernel and a host
ranking of the certificate."""

import numpy as np
import pytest
import torch
from conftest import aperture_grid, rel_l2, surfels

from oracle import surfel as O

pytestmark = pytest.mark.gpu
TOL = 1e-5


def f32(x):
    return torch.as_tensor(np.asarray(x, np.float32)).cuda()


def c64(x):
    return torch.as_tensor(np.asarray(x, np.complex64)).cuda()


def scan_set(bins, sides=(5, 6, 3), seed=0):
    """Square arrays of side x side apertures (25, 36 and 9: ragged 32-aperture blocks)
    with complex Gaussian residuals."""
    rng = np.random.default_rng(seed)
    out = []
    for side in sides:
        p = aperture_grid(side)
        out.append((p, rng.standard_normal((len(p), bins)) + 1j * rng.standard_normal((len(p), bins))))
    return out


@pytest.mark.parametrize("pad,fresnel_n", [(2, 0.0), (2, 1.6), (1, 0.0), (4, 0.0), (8, 0.0)])
@pytest.mark.parametrize("chunk", [8, 24, 0])
def test_batched_scan_matches_the_oracle_per_scan(pad, fresnel_n, chunk):
    from dsplat import FMCWConfig
    from dsplat.kernels import surfel

    cfg = FMCWConfig(pad_factor=pad)
    p = surfels(37, seed=pad)
    scans = scan_set(cfg.num_range_bins, seed=pad)
    corr, norm = surfel.scan_batched(
        cfg, [(f32(x), c64(r)) for x, r in scans], f32(p["centers"]), f32(p["normals"]), fresnel_n, chunk
    )
    assert corr.shape == (3, 37) and norm.shape == (3, 37)
    for i, (x, r) in enumerate(scans):
        wc, wn = O.correlate(x, p["centers"], p["normals"], r, cfg, fresnel_n)
        assert rel_l2(corr[i].cpu().numpy(), wc) < TOL
        assert rel_l2(norm[i].cpu().numpy(), wn) < TOL


def test_batched_scan_equals_the_single_scan_correlate():
    """Same per-block sums as dsplat_surfel_correlate; only the float64 reduction order differs."""
    from dsplat import FMCWConfig
    from dsplat.kernels import surfel

    cfg = FMCWConfig()
    p = surfels(300, seed=9, spread=0.03)
    scans = [(f32(x), c64(r)) for x, r in scan_set(cfg.num_range_bins, sides=(12, 9, 16, 7), seed=9)]
    centers, normals = f32(p["centers"]), f32(p["normals"])
    corr, norm = surfel.scan_batched(cfg, scans, centers, normals, 0.0)
    for i, (x, r) in enumerate(scans):
        c, n = surfel.correlate(cfg, x, centers, normals, r, 0.0)
        assert rel_l2(corr[i].cpu().numpy(), c.cpu().numpy()) < 1e-6
        assert rel_l2(norm[i].cpu().numpy(), n.cpu().numpy()) < 1e-6


def test_batched_scan_is_independent_of_the_chunk():
    from dsplat import FMCWConfig
    from dsplat.kernels import surfel

    cfg = FMCWConfig()
    p = surfels(203, seed=2, spread=0.03)
    scans = [(f32(x), c64(r)) for x, r in scan_set(cfg.num_range_bins, seed=2)]
    results = [surfel.scan_batched(cfg, scans, f32(p["centers"]), f32(p["normals"]), 0.0, k) for k in (8, 64, 0, 4096)]
    for corr, norm in results[1:]:
        assert torch.equal(corr, results[0][0]) and torch.equal(norm, results[0][1])


def test_batched_scan_rejects_more_than_eight_scans_and_allows_no_candidates():
    from dsplat import FMCWConfig
    from dsplat.kernels import surfel

    cfg = FMCWConfig()
    p = surfels(4, seed=1)
    scans = [(f32(x), c64(r)) for x, r in scan_set(cfg.num_range_bins, sides=(2,) * 9)]
    with pytest.raises(RuntimeError):
        surfel.scan_batched(cfg, scans, f32(p["centers"]), f32(p["normals"]), 0.0, 8)
    empty = torch.zeros((0, 3), device="cuda")
    corr, norm = surfel.scan_batched(cfg, scans[:2], empty, empty, 0.0, 8)
    assert corr.shape == (2, 0) and norm.shape == (2, 0)


@pytest.mark.parametrize("mode", ["complex", "real"])
def test_certificate_scan_ranks_like_the_oracle_certificate(mode):
    """Scores equal |s|^2 / (|a|^2 |r|^2) from the oracle's columns with area, and the
    top ids are the stable descending order of those scores."""
    from dsplat import FMCWConfig, FMCWModel, Scan
    from dsplat.certificate import certificate_scan

    cfg = FMCWConfig()
    p = surfels(150, seed=4, spread=0.03)
    areas = p["areas"] * np.linspace(0.5, 2.0, 150)
    scans = scan_set(cfg.num_range_bins, sides=(6, 4), seed=4)
    models = [FMCWModel(Scan(c64(r), f32(x), cfg)) for x, r in scans]
    candidates = dict(centers=f32(p["centers"]), normals=f32(p["normals"]), areas=f32(areas))
    scores, ids = certificate_scan(models, [m.scan.data for m in models], candidates, mode, top=20)
    assert scores.shape == (2, 150) and ids.shape == (2, 20)
    for i, (x, r) in enumerate(scans):
        a = O.columns(x, p["centers"], p["normals"], cfg, 0.0) * np.sqrt(areas)  # [A, K, S]
        a = a.reshape(-1, len(areas))
        s = a.conj().T @ r.reshape(-1)
        s = s.real if mode == "real" else s
        want = np.abs(s) ** 2 / (np.abs(a) ** 2).sum(0) / (np.abs(r) ** 2).sum()
        assert rel_l2(scores[i].cpu().numpy(), want) < 1e-5
        top = scores[i].cpu().numpy()[ids[i].cpu().numpy()]
        assert np.all(np.diff(top) <= 0)
        assert top[-1] >= np.sort(scores[i].cpu().numpy())[-20]
