"""Spectral, search, energy, response and SYRK kernels on the GPU against the oracle."""

import numpy as np
import pytest
import torch
from conftest import rel_l2

from oracle import linalg as L
from oracle import search as Se
from oracle import spectral as Sp

pytestmark = pytest.mark.gpu


def grids():
    from dsplat import SpectralGrid

    return [
        SpectralGrid((32,)),
        SpectralGrid((24, 20), window="hann"),
        SpectralGrid((24, 20), samples=(16, 20), window="hamming"),
        SpectralGrid((16, 12), window="kaiser", beta=6.0),
        SpectralGrid((24, 20), radius=4.0),
        SpectralGrid((8, 6, 4), window="hann"),
    ]


@pytest.mark.parametrize("index", range(6))
def test_spectral_design_matches_the_oracle(index):
    """tests/test_spectral.py:12-33 of the reference: atoms 2e-5, derivatives 1e-4 of max."""
    from dsplat.kernels import spectral

    grid = grids()[index]
    g = torch.Generator().manual_seed(index)
    centers = (torch.rand(7, len(grid.shape), generator=g) * torch.tensor(grid.shape)).cuda()
    bins = grid.bins()
    a, da = spectral.design(grid, centers, bins, derivatives=True)
    ea, eda = Sp.design(grid.shape, grid.samples, grid.window, grid.beta, grid.radius, centers.cpu().numpy(),
                        bins.cpu().numpy())
    if grid.radius:  # the truncated kernel is discontinuous at the support edge
        inside = (np.abs(ea) > 0)[..., None]
        eda, da = eda * inside, da * torch.as_tensor(inside).cuda()
    a, da = a.cpu().numpy(), da.cpu().numpy()
    assert np.abs(a - ea).max() <= 2e-5 * np.abs(ea).max()
    assert np.abs(da - eda).max() <= 1e-4 * np.abs(eda).max()


@pytest.mark.parametrize("index", [0, 1, 3, 5])
def test_spectral_render_jvp_vjp_are_one_linear_map(index):
    """tests/test_spectral.py:36-54: render = A b, JVP = (dA dc) b + A db, adjoint identity,
    autograd equals the explicit VJP."""
    from dsplat.kernels import spectral

    grid = grids()[index]
    g = torch.Generator().manual_seed(10 + index)
    dim = len(grid.shape)
    centers = (torch.rand(3, dim, generator=g) * torch.tensor(grid.shape)).cuda()
    b = torch.randn(3, dtype=torch.complex64, generator=g).cuda()
    bins = grid.bins()
    a, da = spectral.design(grid, centers, bins, derivatives=True)
    render = spectral.render(grid, centers, b, bins)
    assert rel_l2(render.cpu().numpy(), (a @ b).cpu().numpy()) < 2e-5
    dc = torch.randn(3, dim, generator=g).cuda()
    db = torch.randn(3, dtype=torch.complex64, generator=g).cuda()
    jvp = spectral.jvp(grid, centers, b, dc, db, bins)
    assert rel_l2(jvp.cpu().numpy(), ((da * dc[None]).sum(-1) @ b + a @ db).cpu().numpy()) < 1e-4
    up = torch.randn(len(bins), dtype=torch.complex64, generator=g).cuda()
    gc, gb = spectral.vjp(grid, centers, b, up, bins)
    left = float((up.conj() * jvp).real.sum())
    right = float((gc * dc).sum() + (gb.conj() * db).real.sum())
    assert abs(left - right) <= 1e-4 * abs(left) + 1e-6
    c_leaf, b_leaf = centers.clone().requires_grad_(), b.clone().requires_grad_()
    (up.conj() * spectral.render(grid, c_leaf, b_leaf, bins)).real.sum().backward()
    assert torch.equal(c_leaf.grad, gc) and torch.equal(b_leaf.grad, gb)


def test_peaks_match_brute_force_greedy_suppression():
    """tests/test_search.py:19-39: 5000 points, counts 1/10/200, ties by index."""
    from dsplat.certificate import candidate_grid, select_peaks

    grid = candidate_grid([0.0, -1.0], [1.0, 1.0], 0.3)
    assert grid.shape == (5 * 8, 2) and grid.min(0).values.tolist() == [0, -1]
    g = torch.Generator().manual_seed(1)
    centers = torch.rand(5000, 3, generator=g)
    scores = torch.rand(5000, generator=g)
    scores[17] = scores[4000] = scores.max() + 0.1  # an exact tie: the lower index wins
    for count in (1, 10, 200):
        got = select_peaks(scores.cuda(), centers.cuda(), count, 0.1).tolist()
        assert got == Se.peaks(scores.numpy(), centers.numpy(), count, 0.1)
    assert select_peaks(scores.cuda(), centers.cuda(), 3, 0.0).tolist()[:2] == [17, 4000]


@pytest.mark.parametrize("dim,count,separation", [(3, 500, 0.05), (2, 64, 0.2), (1, 40, 0.01), (3, 3000, 0.0)])
def test_peaks_kernel_matches_serial_greedy_selection(dim, count, separation):
    """Many sorted blocks, taken centers beyond the LDS cache (3000 x 3 floats), 1-D/2-D
    points, non-finite scores that are never taken and a -0.0 / 0.0 tie."""
    from dsplat.kernels import search

    g = torch.Generator().manual_seed(dim + count)
    n = 20000
    centers = torch.rand(n, dim, generator=g)
    scores = torch.rand(n, generator=g)
    scores[::97] = float("nan")
    scores[5::101] = float("inf")
    scores[7::103] = -float("inf")
    scores[11], scores[12] = 0.0, -0.0
    scores[13] = scores[14] = 0.5  # an exact tie: the lower index wins
    got = search.peaks(scores.cuda(), centers.cuda(), count, separation).cpu().tolist()
    want = Se.peaks(scores.numpy(), centers.numpy(), count, separation)
    assert got == want
    assert torch.isnan(scores[0])  # the scores were not modified


@pytest.mark.parametrize("dim,clusters,size,count", [(3, 1500, 100, 1000), (2, 400, 60, 500), (5, 300, 40, 200)])
def test_peaks_kernel_skips_clustered_near_duplicates(dim, clusters, size, count):
    """The certificate-scan regime: every strong peak comes with many lower-scored neighbours
    inside the separation, so most sorted candidates are suppressed by a taken peak and the
    prefiltered windows add few survivors; the linear scan for more than 3 coordinates."""
    from dsplat.kernels import search

    g = torch.Generator().manual_seed(dim * clusters + size)
    hubs = torch.rand(clusters, dim, generator=g)
    centers = (hubs[:, None] + 0.004 * (2 * torch.rand(clusters, size, dim, generator=g) - 1)).reshape(-1, dim)
    scores = torch.rand(clusters * size, generator=g)
    scores[::89] = float("nan")
    got = search.peaks(scores.cuda(), centers.cuda(), count, 0.02).cpu().tolist()
    want = Se.peaks(scores.numpy(), centers.numpy(), count, 0.02)
    assert got == want


def test_peaks_kernel_stops_when_every_candidate_is_suppressed():
    from dsplat.kernels import search

    centers = torch.zeros(700, 3)
    centers[:, 0] = torch.arange(700) * 1e-3
    scores = torch.linspace(1, 2, 700)
    got = search.peaks(scores.cuda(), centers.cuda(), 50, 10.0).cpu().tolist()
    assert got == [699]


def test_range_gate_matches_the_oracle_and_the_reference_known_answer():
    from dsplat.kernels import search

    g = torch.Generator().manual_seed(2)
    positions = (torch.rand(40, 3, generator=g) - 0.5) * 0.2
    positions[:, 2] = 0
    points = torch.rand(3000, 3, generator=g) * 0.1 + torch.tensor([-0.05, -0.05, 0.25])
    intervals = (torch.rand(40, 128, generator=g) > 0.7).int()
    spacing = 0.0049965
    got = search.range_gate(points.cuda(), positions.cuda(), intervals.cuda(), spacing, 0.3)
    import math

    required = math.ceil(0.3 * 40 - 1e-7)
    want = Se.range_gate(points.numpy(), positions.numpy(), intervals.numpy(), spacing, required)
    assert (got.cpu().numpy() != want).sum() <= 1  # float32 distance ties at exact half bins


def test_direction_and_trial_match_the_oracle():
    from dsplat.kernels import search

    g = torch.Generator().manual_seed(3)
    for width in (3, 6):
        inv = torch.randn(50, width, width, generator=g)
        inv = inv @ inv.transpose(-1, -2) + 0.1 * torch.eye(width)
        inv[:5] = -inv[:5]  # not descent directions: fall back to -g
        grad = torch.randn(50, width, generator=g)
        normals = torch.nn.functional.normalize(torch.randn(50, 3, generator=g), dim=1)
        mask = torch.ones(width)
        mask[1] = 0
        d = search.direction(inv.cuda(), grad.cuda(), normals.cuda(), mask.cuda(), 0.7)
        want = Se.direction(inv.numpy(), grad.numpy(), normals.numpy(), mask.numpy(), 0.7)
        assert np.abs(d.cpu().numpy() - want).max() < 1e-5
        controls = torch.tensor([[0.5, 0.5, 0.5, 0.1, 0.1, 0.1], [1, 0, 1, 1, 1, 1],
                                 [-0.1, -0.1, -0.1, 0, 0, 0], [0.1, 0.1, 0.1, 0, 0, 0]], dtype=torch.float32)
        centers = torch.rand(50, 3, generator=g) * 0.2 - 0.1
        oc, on = search.trial(centers.cuda(), normals.cuda(), d, controls.cuda(), 0.5)
        wc, wn = Se.trial(centers.numpy(), normals.numpy(), d.cpu().numpy(), controls.numpy(), 0.5)
        assert np.abs(oc.cpu().numpy() - wc).max() < 1e-6 and np.abs(on.cpu().numpy() - wn).max() < 1e-6


@pytest.mark.parametrize("count", [1, 255, 256, 100003, 3_000_000])
def test_energy_matches_a_float64_sum(count):
    from dsplat.kernels.reduce import energy

    g = torch.Generator().manual_seed(count)
    x = torch.randn(count, dtype=torch.complex64, generator=g)
    y = torch.randn(count, dtype=torch.complex64, generator=g)
    got = float(energy(x.cuda(), y.cuda()))
    want = L.energy(x.numpy(), y.numpy())
    assert abs(got - want) <= 1e-6 * want


def test_range_response_matches_the_oracle():
    from dsplat import FMCWConfig
    from dsplat.kernels.response import range_response

    cfg = FMCWConfig()
    got = range_response(cfg, 55.3, 78.9, 401, "cuda").cpu().numpy()
    want = L.range_response(55.3, 78.9, 401, 128, 128, 256)
    assert np.abs(got - want).max() < 1e-13


SYRK_BARS = {"fp32": 2e-6, "bf16x3": 6e-5, "fp16x3": 2e-6}


@pytest.mark.parametrize("mode", ["fp32", "bf16x3", "fp16x3"])
@pytest.mark.parametrize("m,n", [(5, 7), (256, 64), (1000, 130), (4096, 513), (20000, 300)])
def test_syrk_updates_the_upper_triangle_only(m, n, mode):
    """Every mode against the float64 product: the lower triangle is untouched, the upper
    one gets rows^T rows within the mode's bar (relative to the largest entry)."""
    from dsplat.kernels.blas import syrk

    g = torch.Generator().manual_seed(m + n)
    rows = torch.randn(m, n, generator=g)
    out = torch.randn(n, n, generator=g)
    before = out.clone()
    got = syrk(rows.cuda(), out.cuda(), mode).cpu()
    full = before.double() + rows.double().T @ rows.double()
    upper = torch.triu(torch.ones(n, n, dtype=torch.bool))
    assert torch.equal(got[~upper], before[~upper])
    err = (got.double() - full)[upper].abs().max() / full[upper].abs().max()
    assert err < SYRK_BARS[mode]


@pytest.mark.parametrize("mode", ["bf16x3", "fp16x3"])
def test_split_syrk_keeps_columns_of_very_different_scale(mode):
    """Columns scaled from 1e-12 to 1e12, and an all-zero column: each Gram entry is
    within the bar of its own scale sqrt(G_ii G_jj) (FP16 needs the per-column scaling)."""
    from dsplat.kernels.blas import syrk

    g = torch.Generator().manual_seed(7)
    n = 160
    rows = torch.randn(3000, n, generator=g) * torch.logspace(-12, 12, n)
    rows[:, 17] = 0
    got = syrk(rows.cuda(), torch.zeros(n, n, device="cuda"), mode).cpu().double()
    want = rows.double().T @ rows.double()
    scale = want.diagonal().sqrt()
    bound = scale[:, None] * scale[None, :]
    upper = torch.triu(torch.ones(n, n, dtype=torch.bool)) & (bound > 0)
    assert ((got - want).abs()[upper] / bound[upper]).max() < SYRK_BARS[mode]
    assert torch.all(got[17] == 0) and torch.all(got[:, 17] == 0)


def test_syrk_is_deterministic():
    from dsplat.kernels.blas import syrk

    rows = torch.randn(9000, 700, generator=torch.Generator().manual_seed(3)).cuda()
    first = syrk(rows, torch.zeros(700, 700, device="cuda"), "fp16x3")
    for _ in range(2):
        assert torch.equal(syrk(rows, torch.zeros(700, 700, device="cuda"), "fp16x3"), first)
