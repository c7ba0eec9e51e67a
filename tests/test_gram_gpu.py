"""Fused split-precision Gram against dense float64 products of the design, and the oracle."""

import numpy as np
import pytest
import torch
from conftest import aperture_grid, rel_l2, surfels

from oracle import surfel as O

pytestmark = pytest.mark.gpu

# Frobenius relative error bars of each mode against float64 (measured values are printed)
BARS = {"fp32": 2e-6, "bf16x3": 5e-5, "fp16x3": 2e-6}


def scene(count, side, seed):
    p = surfels(count, seed=seed, spread=0.02)
    return (torch.as_tensor(aperture_grid(side)).cuda(), torch.as_tensor(p["centers"]).cuda(),
            torch.as_tensor(p["normals"]).cuda(), torch.as_tensor(p["areas"]).cuda() * torch.linspace(0.5, 2, count).cuda())


@pytest.mark.parametrize("mode", ["fp32", "bf16x3", "fp16x3"])
@pytest.mark.parametrize("count", [37, 300])
def test_gram_matches_dense_float64_products(mode, count):
    from dsplat import FMCWConfig
    from dsplat.kernels import blas, surfel

    cfg = FMCWConfig()
    positions, centers, normals, areas = scene(count, 8, count)
    a = surfel.design(cfg, positions, centers, normals, areas, 0.0, False).to(torch.complex128)
    m = a.shape[0]
    dense = (a.H @ a) / m
    g = blas.gram(cfg, positions, centers, normals, areas, mode, real_only=False, scale=1.0 / m)
    err = float((g.to(torch.complex128) - dense).norm() / dense.norm())
    print(f"gram {mode} S={count}: rel Frobenius {err:.3e}", flush=True)
    assert err < BARS[mode]
    assert torch.equal(g, g.conj().T)  # exactly Hermitian by construction
    real = blas.gram(cfg, positions, centers, normals, areas, mode, real_only=True, scale=1.0 / m)
    assert float((real.double() - dense.real).norm() / dense.real.norm()) < BARS[mode]


def test_gram_matches_the_oracle_design():
    from dsplat import FMCWConfig
    from dsplat.kernels import blas

    cfg = FMCWConfig(pad_factor=4)
    positions, centers, normals, areas = scene(20, 6, 3)
    a = O.design(positions.cpu().numpy(), centers.cpu().numpy(), normals.cpu().numpy(), areas.cpu().numpy(), cfg)
    dense = a.conj().T @ a
    for mode in ("fp32", "fp16x3"):
        g = blas.gram(cfg, positions, centers, normals, areas, mode).cpu().numpy()
        assert rel_l2(g, dense) < BARS[mode]


# (G, coefficients) bars at 2^21 rows, about twice to three times the measured errors:
# fp32 2.2e-5 / 1.2e-3, bf16x3 3.0e-6 / 1.8e-4, fp16x3 1.6e-6 / 7.0e-5
FULL_BARS = {"fp32": (5e-5, 3e-3), "bf16x3": (1e-5, 5e-4), "fp16x3": (5e-6, 2e-4)}


def test_full_size_gram_and_solve_against_float64():
    """The Gram as benchmarked: 128^2 apertures x 128 bins (2^21 rows) x the 2048-surfel
    ellipsoid dictionary, seed-3 right-hand side. Each mode's G = A^H A / M against float64
    products of the same float32 atoms, A^H r by the correlation kernel, and the coefficients
    of a ridge-1e-2 solve (ridge relative to the mean diagonal) against the float64 solve."""
    from dsplat import workloads
    from dsplat.kernels import blas, surfel
    from dsplat.varpro import solve

    work = workloads.ellipsoid_scene()
    cfg = work["config"]
    pos, c, n, areas = (work[k].cuda() for k in ("positions", "centers", "normals", "areas"))
    A, K, S = len(pos), cfg.num_range_bins, len(c)
    M = A * K
    residual = workloads.complex_normal((M,), torch.Generator().manual_seed(3)).cuda()
    g64 = torch.zeros((S, S), dtype=torch.complex128, device="cuda")
    q64 = torch.zeros(S, dtype=torch.complex128, device="cuda")
    for start in range(0, A, 256):
        a = surfel.design(cfg, pos[start : start + 256], c, n, areas, 0.0, False).to(torch.complex128)
        g64 += a.H @ a
        q64 += a.H @ residual[start * K : (start + 256) * K].to(torch.complex128)
        del a
    g64, q64 = g64 / M, q64 / M
    corr, _ = surfel.correlate(cfg, pos, c, n, residual.reshape(A, K), 0.0)
    q = (corr * areas.sqrt()).to(torch.complex128) / M
    q_err = float((q - q64).norm() / q64.norm())
    print(f"rhs: rel {q_err:.3e}", flush=True)
    assert q_err < 1e-5
    eye = torch.eye(S, dtype=torch.complex128, device="cuda")
    ridge = 1e-2 * float(g64.diagonal().real.mean())
    b64 = solve(g64 + ridge * eye, q64)
    for mode in ("fp32", "bf16x3", "fp16x3"):
        g = blas.gram(cfg, pos, c, n, areas, mode, scale=1.0 / M).to(torch.complex128)
        err = float((g - g64).norm() / g64.norm())
        b = solve(g + ridge * eye, q)
        b_err = float((b - b64).norm() / b64.norm())
        print(f"gram {mode}: rel Frobenius {err:.3e}, coefficients {b_err:.3e}", flush=True)
        assert err < FULL_BARS[mode][0]
        assert b_err < FULL_BARS[mode][1]
