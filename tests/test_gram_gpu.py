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
