"""The solver layer on the GPU kernels, after the reference's test strategy
(tests/test_fmcw.py, test_solver.py, test_search.py): dense-algebra equivalences,
one-column updates equal full refits, monotone objectives, and recovery of known scenes."""

import itertools
import math

import pytest
import torch

pytestmark = pytest.mark.gpu


def close(actual, expected, tolerance=2e-5):
    expected = expected.to(actual.device, actual.dtype)
    atol = tolerance * float(expected.abs().max())
    torch.testing.assert_close(actual, expected, rtol=0, atol=atol)


def aperture_grid(side=8, pitch=0.004):
    t = (torch.arange(side, dtype=torch.float32) - (side - 1) / 2) * pitch
    x, y = torch.meshgrid(t, t, indexing="ij")
    return torch.stack([x.flatten(), y.flatten(), torch.zeros(side * side)], -1).cuda()


def surfels(count, seed=0, spread=0.01):
    generator = torch.Generator().manual_seed(seed)
    centers = torch.tensor([0, 0, 0.3]) + spread * (2 * torch.rand(count, 3, generator=generator) - 1)
    normals = torch.nn.functional.normalize(
        torch.tensor([0.0, 0, -1]) + 0.3 * torch.randn(count, 3, generator=generator), dim=1
    )
    areas = torch.full((count,), 1e-4)
    return {k: v.cuda() for k, v in dict(centers=centers, normals=normals, areas=areas).items()}


def fmcw_specs(trainable_normals=True):
    from dsplat import ParameterConfig

    return dict(
        centers=ParameterConfig(scale=0.0005, lower=(-0.02, -0.02, 0.28), upper=(0.02, 0.02, 0.32)),
        normals=ParameterConfig(trainable=trainable_normals, scale=0.1, normalize=True),
        areas=ParameterConfig(trainable=False),
    )


@pytest.fixture
def fmcw():
    from dsplat import FMCWConfig, FMCWModel, Scan

    config = FMCWConfig()
    positions = aperture_grid()
    params = surfels(5)
    coefficients = torch.linspace(0.5, 1.5, 5, device="cuda")
    empty = Scan(torch.zeros((len(positions), config.num_range_bins), dtype=torch.complex64, device="cuda"),
                 positions, config)
    data = FMCWModel(empty).render(params, coefficients)
    return FMCWModel(Scan(data, positions, config)), params, coefficients


@pytest.fixture
def spectral():
    from dsplat import SpectralGrid, SpectralModel

    grid = SpectralGrid((24, 20))
    centers = torch.tensor([[3.3, 4.1], [12.0, 15.0], [18.4, 7.9]], device="cuda")
    coefficients = torch.tensor([1 + 0.5j, -0.7 + 0.2j, 0.4 - 0.9j], device="cuda")
    return SpectralModel(grid), {"centers": centers}, coefficients


# --- measurement model -----------------------------------------------------------------


@pytest.mark.parametrize("pad", [2, 4])
def test_row_basis_preserves_the_normal_equations(pad):
    from dsplat import FMCWConfig, FMCWModel, Scan
    from dsplat.fmcw import response_basis

    config = FMCWConfig(pad_factor=pad)
    data = torch.zeros((36, config.num_range_bins), dtype=torch.complex64, device="cuda")
    model = FMCWModel(Scan(data, aperture_grid(6), config))
    a = model.design(surfels(6), slice(0, model.size))
    spacing = config.range_bin_spacing_m
    reduced = response_basis(config, 0.28 / spacing, 0.33 / spacing, "cuda").reduce(a)
    close(reduced.H @ reduced, a.H @ a, 1e-4)


def test_range_gate_keeps_points_at_measured_ranges(fmcw):
    model, params, _ = fmcw
    away = params["centers"] + torch.tensor([0, 0, 0.05], device="cuda")
    keep = model.range_gate(torch.cat((params["centers"], away)), model.scan.data, 0.1, margin_cells=1.0,
                            min_fraction=0.9)
    assert keep[:5].all() and not keep[5:].any()


def test_fft_range_table_approximates_the_exact_proposal_scores(fmcw):
    model, params, _ = fmcw
    residual = model.scan.data
    bank = torch.nn.functional.normalize(torch.randn(6, 3, device="cuda", generator=torch.Generator("cuda").manual_seed(0)), dim=1)
    exact, norm = model.correlate_orientations(params, residual, bank)
    table = model.range_table(residual, 8)
    approximate, approximate_norm = model.correlate_orientations(params, residual, bank, table)
    close(approximate, exact, 3e-2)
    close(approximate_norm, norm, 3e-2)


def test_render_matches_independent_adc_synthesis():
    """tests/test_surfel.py:12-20 of the reference, at this build's 1e-5 bar."""
    import numpy as np

    from dsplat import FMCWConfig, surfel
    from oracle import surfel as O

    config = FMCWConfig(pad_factor=4)
    positions = aperture_grid(6)
    params = surfels(4)
    reflectivity = torch.rand(4, device="cuda", generator=torch.Generator("cuda").manual_seed(1)) + 0.5
    rendered = surfel.render(config, positions, params, reflectivity).cpu().numpy()
    amplitudes = (reflectivity * params["areas"].sqrt()).cpu().numpy()
    synth = O.synthesize(positions.cpu().numpy(), params["centers"].cpu().numpy(), params["normals"].cpu().numpy(),
                         amplitudes, config)
    assert np.linalg.norm(rendered - synth) / np.linalg.norm(synth) < 1e-5


# --- varpro ----------------------------------------------------------------------------


def test_streamed_normal_equations_match_dense_products(fmcw):
    from dsplat import DSFWConfig
    from dsplat.varpro import normal_equations

    model, params, _ = fmcw
    model = model.prepare(fmcw_specs())
    y = model.scan.data.flatten()
    a = model.design(params, slice(0, model.size)).to(torch.complex128)
    m = model.size
    gram, rhs = normal_equations(model, params, model.scan.data, DSFWConfig(row_chunk=2 * model.shape[1]))
    close(gram, (a.H @ a).real / m, 1e-4)
    close(rhs, (a.H @ y.to(torch.complex128)).real / m, 1e-4)
    gram, rhs = normal_equations(model, params, model.scan.data, DSFWConfig(coefficients="complex"))
    close(gram, a.H @ a / m, 1e-4)
    close(rhs, a.H @ y.to(torch.complex128) / m, 1e-4)


def test_varpro_minimizes_the_ridge_objective(spectral):
    from dsplat import DSFWConfig
    from dsplat.varpro import varpro

    model, params, b = spectral
    target = model.render(params, b)
    trial = {"centers": params["centers"] + 0.3}
    config = DSFWConfig(coefficients="complex", ridge=1e-3)
    coefficients, prediction, loss, _, _ = varpro(model, trial, target, config)
    a = model.design(trial, slice(0, model.size)).to(torch.complex128)
    m = model.size
    h = a.H @ a / m + 1e-3 * torch.eye(3, device="cuda")
    expected = torch.linalg.solve(h, a.H @ target.flatten().to(torch.complex128) / m)
    close(coefficients, expected, 1e-4)
    rendered = (prediction - target).abs().square().mean() + 1e-3 * coefficients.abs().square().sum()
    assert abs(loss - float(rendered)) <= 1e-5 * float(rendered)


def test_one_column_exchange_equals_a_full_refit(fmcw):
    from dsplat import DSFWConfig, initialize
    from dsplat.varpro import exchange_varpro, varpro

    model, params, _ = fmcw
    config = DSFWConfig()
    state = initialize(model, model.scan.data, {**params, "centers": params["centers"] + 1e-4}, config,
                       parameters=fmcw_specs())
    proposal = {k: v.clone() for k, v in state.params.items()}
    proposal["centers"][2] = params["centers"][2]
    fast = exchange_varpro(state, proposal, 2, config)
    full = varpro(model, proposal, model.scan.data, config)
    close(fast.coefficients, full.coefficients, 1e-3)
    assert abs(fast.objective - full.objective) <= 1e-3 * full.objective


def test_pooled_statistics_match_a_full_stream(fmcw):
    from dsplat import DSFWConfig, initialize
    from dsplat.replacement import pooled_statistics
    from dsplat.varpro import normal_equations

    model, params, _ = fmcw
    config = DSFWConfig()
    state = initialize(model, model.scan.data, surfels(5, seed=2), config, parameters=fmcw_specs())
    pool = {k: torch.cat((v, params[k])) for k, v in state.params.items()}
    gram, rhs = pooled_statistics(state, pool, config)
    expected_gram, expected_rhs = normal_equations(state.model, pool, model.scan.data, config)
    close(gram, expected_gram, 1e-4)
    close(rhs, expected_rhs, 1e-4)


# --- replacement, refinement, search -----------------------------------------------------


def test_merged_replacement_restores_displaced_surfels(fmcw):
    from dsplat import DSFWConfig, initialize
    from dsplat.replacement import merge_support

    model, params, _ = fmcw
    config = DSFWConfig(batch=5, screen_apertures=0)
    state = initialize(model, model.scan.data, surfels(5, seed=7), config, candidates=params, parameters=fmcw_specs())
    after, replaced = merge_support(state, config, state.candidates)
    assert replaced > 0 and after.objective < state.objective and len(after.coefficients) == 5


def test_utility_replacement_recovers_a_misplaced_tone(spectral):
    from dsplat import DSFWConfig, ParameterConfig, initialize
    from dsplat.replacement import replace_one

    model, params, b = spectral
    target = model.render(params, b)
    wrong = params["centers"].clone()
    wrong[1] = torch.tensor([20.0, 2.0], device="cuda")
    config = DSFWConfig(coefficients="complex", replacement="utility")
    state = initialize(model, target, {"centers": wrong}, config, candidates={"centers": model.grid.bins()},
                       parameters={"centers": ParameterConfig(period=model.shape)})
    after, replaced = replace_one(state, config, state.candidates)
    assert replaced == 1 and after.objective < 0.1 * state.objective


def test_joint_system_is_symmetric_and_pcg_matches_a_dense_solve(spectral):
    from dsplat import DSFWConfig, ParameterConfig
    from dsplat.refinement import joint_system
    from dsplat.varpro import pcg

    model, params, b = spectral
    target = model.render(params, b)
    trial = {"centers": params["centers"] + 0.1}
    config = DSFWConfig(coefficients="complex", ridge=1e-4)
    matvec, gradient, diagonal, _ = joint_system(model, trial, b, target, {"centers": ParameterConfig()}, config, 0)
    columns = torch.eye(len(gradient), device="cuda")
    dense = torch.stack([matvec(column) for column in columns], 1).double()
    close(dense, dense.T, 1e-4)
    assert torch.linalg.eigvalsh((dense + dense.T) / 2)[0] > 0
    solution = pcg(matvec, -gradient, diagonal, iterations=200, tolerance=1e-7)
    close(solution, torch.linalg.solve(dense, -gradient.double()), 1e-3)


def test_fused_pcg_matches_a_dense_solve_and_the_tensor_iteration(spectral):
    """The one-kernel iteration (float64 sums) converges to the dense solve and follows the
    tensor-op iteration to rounding at the solver's 16 iterations."""
    from dsplat import DSFWConfig, ParameterConfig
    from dsplat.refinement import joint_system
    from dsplat.varpro import pcg

    model, params, b = spectral
    target = model.render(params, b)
    trial = {"centers": params["centers"] + 0.1}
    config = DSFWConfig(coefficients="complex", ridge=1e-4)
    matvec, gradient, diagonal, _ = joint_system(model, trial, b, target, {"centers": ParameterConfig()}, config, 0)
    dense = torch.stack([matvec(column) for column in torch.eye(len(gradient), device="cuda")], 1).double()
    solution = pcg(matvec, -gradient, diagonal, iterations=200, tolerance=1e-7, fused=True)
    close(solution, torch.linalg.solve(dense, -gradient.double()), 1e-3)
    close(pcg(matvec, -gradient, diagonal, iterations=16, fused=True), pcg(matvec, -gradient, diagonal, iterations=16),
          1e-4)  # fmt: skip


@pytest.mark.parametrize("case", ["indefinite", "converges"])
def test_fused_pcg_stops_where_the_tensor_iteration_stops(case):
    """Non-positive curvature stops before the update and a small residual after it, as in
    the tensor-op iteration; later iterations change nothing."""
    from dsplat.varpro import pcg

    g = torch.Generator().manual_seed(5)
    x = torch.randn(400, 120, generator=g)
    h = (x.T @ x / 400 + 0.05 * torch.eye(120)).cuda()
    if case == "indefinite":
        h[7, 7] = -30.0
    rhs = torch.randn(120, generator=g).cuda()
    tolerance = 1e-12 if case == "indefinite" else 1e-2
    want = pcg(lambda v: h @ v, rhs, h.diag().abs(), iterations=60, tolerance=tolerance)
    got = pcg(lambda v: h @ v, rhs, h.diag().abs(), iterations=60, tolerance=tolerance, fused=True)
    close(got, want, 1e-4)
    longer = pcg(lambda v: h @ v, rhs, h.diag().abs(), iterations=90, tolerance=tolerance, fused=True)
    assert torch.equal(longer, got)


@pytest.mark.parametrize("iterations", [2, 5, 16])
def test_pcg_graph_replays_equal_the_eager_iterations(fmcw, iterations):
    """The FMCW joint system (JVP and VJP kernels, tangent projections, area chain): one
    iteration captured in a HIP graph and replayed gives the eager fused iterations bit for bit,
    run after run, with eager work and synchronizations between the runs."""
    from dsplat import DSFWConfig
    from dsplat.refinement import joint_system
    from dsplat.varpro import pcg

    model, params, b = fmcw
    trial = dict(params, centers=params["centers"] + 2e-4)
    matvec, gradient, diagonal, _ = joint_system(model, trial, b, model.scan.data, fmcw_specs(), DSFWConfig(), 0)
    eager = pcg(matvec, -gradient, diagonal, iterations=iterations, fused=True)
    assert torch.isfinite(eager).all() and eager.abs().max() > 0
    for _ in range(3):
        graphed = pcg(matvec, -gradient, diagonal, iterations=iterations, fused=True, graph=True)
        torch.cuda.synchronize()
        assert torch.equal(graphed, eager)
        pcg(matvec, -gradient, diagonal, iterations=3)  # eager kernels and syncs between graph runs
        torch.cuda.synchronize()


@pytest.mark.parametrize("coefficients,normals,areas", [("real", True, False), ("complex", True, True),
                                                      ("real", False, True)])
def test_fused_joint_maps_match_the_tensor_operator(fmcw, coefficients, normals, areas):
    """The joint operator as unpack kernel, JVP, VJP and pack kernel against its tensor form
    (refinement.py:136-174) on the same system: the same image to float32 rounding, for real
    and complex coefficients, unit normals trained or frozen, areas trained or frozen."""
    from dsplat import DSFWConfig, ParameterConfig
    from dsplat.refinement import joint_system

    model, params, b = fmcw
    trial = dict(params, centers=params["centers"] + 2e-4)
    specs = fmcw_specs(trainable_normals=normals)
    if areas:
        specs["areas"] = ParameterConfig(scale=1e-5)
    config = DSFWConfig(coefficients=coefficients)
    matvec, gradient, _, _ = joint_system(model, trial, b, model.scan.data, specs, config, 0)
    system = matvec.__self__
    assert system._fused is not None
    g = torch.Generator(device="cuda").manual_seed(3)
    for _ in range(3):
        vector = torch.randn(len(gradient), device="cuda", generator=g)
        fused = matvec(vector)
        maps, system._fused = system._fused, None
        tensor = matvec(vector)
        system._fused = maps
        err = float((fused - tensor).norm() / tensor.norm())
        print(f"{coefficients} normals={normals} areas={areas}: fused vs tensor rel L2 {err:.2e}")
        assert err < 1e-5


def test_dense_lm_moves_tones_back(spectral):
    from dsplat import DSFWConfig, ParameterConfig, initialize
    from dsplat.refinement import dense_slide

    model, params, b = spectral
    target = model.render(params, b)
    config = DSFWConfig(coefficients="complex", lm_solver="dense", lm_steps=5, lm_rows=model.size, screen_apertures=0)
    specs = {"centers": ParameterConfig(period=model.shape, scale=0.5)}
    state = initialize(model, target, {"centers": params["centers"] + 0.15}, config, parameters=specs)
    after, accepted = dense_slide(state, config)
    assert accepted > 0 and after.objective < 1e-3 * state.objective


def test_screened_steps_never_raise_the_objective(fmcw):
    from dsplat import DSFWConfig, initialize, step

    model, params, _ = fmcw
    config = DSFWConfig(batch=5, lm_gain_ratio=0, joint_apertures=16, screen_apertures=16)
    state = initialize(model, model.scan.data, surfels(5, seed=9), config, candidates=params, parameters=fmcw_specs())
    for _ in range(4):
        state = step(state, config)
    objectives = [entry["objective"] for entry in state.history]
    assert all(b <= a for a, b in itertools.pairwise(objectives)) and objectives[-1] < objectives[0]


def test_certificate_gradient_matches_finite_differences(fmcw):
    from dsplat.certificate import certificate

    model, params, _ = fmcw
    residual = model.scan.data
    probe = {**params, "centers": params["centers"] + 2e-4}
    _, gradient = certificate(model, probe, residual, "real", gradient=True, normals=True)
    for axis in range(3):
        shift = torch.zeros_like(probe["centers"])
        shift[:, axis] = 1e-5
        plus = certificate(model, {**probe, "centers": probe["centers"] + shift}, residual, "real")
        minus = certificate(model, {**probe, "centers": probe["centers"] - shift}, residual, "real")
        numeric = (plus - minus) / 2e-5
        torch.testing.assert_close(gradient[:, axis], numeric, rtol=2e-2, atol=1e-2 * float(gradient.abs().max()))


def test_continuous_refinement_reaches_a_single_reflector(fmcw):
    from dsplat.certificate import CertificateSearchConfig, search_certificate

    model, params, coefficients = fmcw
    one = {k: v[:1] for k, v in params.items()}
    residual = model.render(one, coefficients[:1])
    start = {k: v.repeat(8, *([1] * (v.ndim - 1))) for k, v in one.items()}
    start["centers"] = start["centers"] + 4e-4 * torch.randn(start["centers"].shape, device="cuda",
                                                             generator=torch.Generator("cuda").manual_seed(0))
    specs = fmcw_specs()
    params_out, scores = search_certificate(
        model, residual, start, specs["centers"], CertificateSearchConfig(peaks=8, separation=0), "real",
        normal_spec=specs["normals"],
    )
    assert float((params_out["centers"][0] - one["centers"][0]).norm()) < 5e-5
    assert scores[0] > 0.99


def test_spatial_search_proposes_the_true_reflector():
    from dsplat import DSFWConfig, FMCWConfig, FMCWModel, Scan, SpatialSearchConfig, initialize
    from dsplat.certificate import CertificateSearchConfig
    from dsplat.spatial import propose

    config = FMCWConfig()
    positions = aperture_grid(8, 0.02)
    truth = surfels(1, seed=5)
    empty = Scan(torch.zeros((64, config.num_range_bins), dtype=torch.complex64, device="cuda"), positions, config)
    data = FMCWModel(empty).render(truth, torch.ones(1, device="cuda"))
    model = FMCWModel(Scan(data, positions, config))
    start = {**truth, "normals": torch.tensor([[0.0, 0.0, 1.0]], device="cuda")}
    dsfw_config = DSFWConfig(batch=1)
    state = initialize(model, data, start, dsfw_config, parameters=fmcw_specs())
    search = SpatialSearchConfig(proposal_apertures=64, refine_apertures=0, cache_peaks=0,
                                 continuous=CertificateSearchConfig(peaks=16))
    candidates, scores, _ = propose(state, dsfw_config, search)
    assert float((candidates["centers"][0] - truth["centers"][0]).norm()) < 1e-4
    assert math.isfinite(float(scores[0]))


def test_geometry_step_backs_off_while_it_gains_little(spectral):
    from dsplat import DSFWConfig, ParameterConfig, initialize, step

    model, params, b = spectral
    target = model.render(params, b)
    wrong = params["centers"].clone()
    wrong[1] = torch.tensor([20.0, 2.0], device="cuda")
    config = DSFWConfig(coefficients="complex", replacement="utility", lm_solver="dense", lm_rows=model.size,
                        lm_gain_ratio=1e9)
    specs = {"centers": ParameterConfig(period=model.shape, scale=0.5)}
    state = initialize(model, target, {"centers": wrong}, config, candidates={"centers": model.grid.bins()},
                       parameters=specs)
    state = step(state, config)
    assert state.lm_wait == 1 and state.lm_skip == 1
    state = step(state, config)
    assert state.history[-1]["lm_accepted"] == 0 and state.lm_skip == 0
