"""Host chains above the kernels on the GPU: the surfel area chain (w = b sqrt(area)) against
float64 differences of the oracle and the adjoint identity, and the spectral model's FFT
correlation of atoms on every bin against atom-by-atom sums."""

import numpy as np
import pytest
import torch
from conftest import aperture_grid, rel_l2, surfels

from oracle import spectral as OS
from oracle import surfel as O

pytestmark = pytest.mark.gpu


def cuda(x, dtype=torch.float32):
    return torch.as_tensor(np.asarray(x)).to("cuda", dtype)


@pytest.mark.parametrize("fresnel_n", [0.0, 1.6])
def test_area_chain_matches_float64_differences_and_the_adjoint(fresnel_n):
    """d render / d area along a direction equals the oracle's central difference in float64
    (render is linear in w = b sqrt(area), so only sqrt's curvature enters, at 1e-8 for the
    step taken), and the area gradient of the VJP satisfies Re<u, J dA> = <g_area, dA>."""
    from dsplat import FMCWConfig
    from dsplat import surfel as S

    cfg = FMCWConfig()
    positions = aperture_grid(8)
    p = surfels(24, seed=11)
    rng = np.random.default_rng(11)
    areas = p["areas"] * rng.uniform(0.5, 2.0, len(p["areas"]))
    b = rng.standard_normal(24) + 1j * rng.standard_normal(24)
    d_area = areas * rng.uniform(-0.5, 0.5, 24)
    params = dict(centers=cuda(p["centers"]), normals=cuda(p["normals"]), areas=cuda(areas))
    tangent = dict(centers=torch.zeros_like(params["centers"]), normals=torch.zeros_like(params["normals"]),
                   areas=cuda(d_area))  # fmt: skip
    coefficients = cuda(b, torch.complex64)
    got = S.jvp(cfg, cuda(positions), params, coefficients, tangent, torch.zeros_like(coefficients), fresnel_n)

    step = 1e-4
    plus = O.render(positions, p["centers"], p["normals"], b * np.sqrt(areas + step * d_area), cfg, fresnel_n)
    minus = O.render(positions, p["centers"], p["normals"], b * np.sqrt(areas - step * d_area), cfg, fresnel_n)
    want = (plus - minus) / (2 * step)
    assert rel_l2(got.cpu().numpy(), want) < 1e-5

    u = rng.standard_normal(want.shape) + 1j * rng.standard_normal(want.shape)
    grads, _ = S.vjp(cfg, cuda(positions), params, coefficients, cuda(u, torch.complex64), fresnel_n)
    left = float(np.real(np.vdot(u, want)))
    right = float(np.dot(grads["areas"].double().cpu().numpy(), d_area))
    assert abs(left - right) <= 1e-5 * np.abs(u).sum() * np.abs(want).max()


@pytest.mark.parametrize("shape,window", [((32,), "hann"), ((12, 10), "rect"), ((8, 6, 5), "hamming")])
def test_spectral_correlation_of_atoms_on_every_bin(shape, window):
    """Atoms centred on every bin take the FFT path (a circular cross-correlation with the
    atom at the origin); it equals the float64 oracle's A^H r and the atom-by-atom VJP path
    (the same atoms in reversed order, which is not the lattice), and every atom's energy is
    the oracle column norm."""
    from dsplat import SpectralGrid, SpectralModel

    grid = SpectralGrid(shape, window=window)
    model = SpectralModel(grid)
    rng = np.random.default_rng(len(shape))
    residual = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    lattice = grid.bins("cuda")
    corr, norm = model.correlate({"centers": lattice}, cuda(residual, torch.complex64))

    bins = OS.grid_bins(shape)
    a, _ = OS.design(shape, shape, window, 14.0, 0.0, bins, bins)
    want = a.conj().T @ residual.reshape(-1)
    assert rel_l2(corr.cpu().numpy(), want) < 1e-5
    np.testing.assert_allclose(norm.double().cpu().numpy(), (np.abs(a) ** 2).sum(0), rtol=1e-5)

    reversed_corr, reversed_norm = model.correlate({"centers": lattice.flip(0)}, cuda(residual, torch.complex64))
    assert rel_l2(reversed_corr.flip(0).cpu().numpy(), corr.cpu().numpy()) < 1e-5
    np.testing.assert_allclose(reversed_norm.flip(0).cpu().numpy(), norm.cpu().numpy(), rtol=1e-5)
