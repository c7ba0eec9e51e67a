"""Shared fixtures. GPU tests carry @pytest.mark.gpu; everything else runs on the CPU."""

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
for path in (ROOT, ROOT / "dirichlet-splatting_amd"):
    if str(path) not in sys.path:
        sys.path.insert(0, str(path))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def pytest_collection_modifyitems(config, items):
    import torch

    if torch.cuda.is_available():
        return
    skip = pytest.mark.skip(reason="no GPU visible; run with -m gpu on an MI355X")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


def rel_l2(actual, expected):
    """||actual - expected|| / ||expected|| in float64."""
    a = np.asarray(actual, dtype=np.complex128 if np.iscomplexobj(actual) else np.float64)
    e = np.asarray(expected, dtype=np.complex128 if np.iscomplexobj(expected) else np.float64)
    return float(np.linalg.norm((a - e).ravel()) / max(np.linalg.norm(e.ravel()), 1e-300))


def close(actual, expected, tolerance=2e-5):
    """Agreement within tolerance times the largest expected magnitude (the reference's
    tests/conftest.py:20-24 convention)."""
    a, e = np.asarray(actual), np.asarray(expected)
    atol = tolerance * float(np.abs(e).max())
    np.testing.assert_allclose(a, e, rtol=0, atol=atol)


def aperture_grid(side=8, pitch=0.004):
    """side x side apertures in z = 0 (reference tests/conftest.py:27-30), float32 numpy."""
    t = (np.arange(side, dtype=np.float32) - (side - 1) / 2) * np.float32(pitch)
    x, y = np.meshgrid(t, t, indexing="ij")
    return np.stack([x.ravel(), y.ravel(), np.zeros(side * side, np.float32)], -1).astype(np.float32)


def surfels(count, seed=0, spread=0.01):
    """Surfels near (0, 0, 0.3) facing the apertures (reference tests/conftest.py:33-43)."""
    import torch

    generator = torch.Generator().manual_seed(seed)
    centers = torch.tensor([0, 0, 0.3]) + spread * (2 * torch.rand(count, 3, generator=generator) - 1)
    normals = torch.nn.functional.normalize(
        torch.tensor([0.0, 0, -1]) + 0.3 * torch.randn(count, 3, generator=generator), dim=1
    )
    areas = torch.full((count,), 1e-4)
    return {k: v.numpy() for k, v in dict(centers=centers, normals=normals, areas=areas).items()}


@pytest.fixture
def chirp():
    from dsplat import FMCWConfig

    return FMCWConfig()
