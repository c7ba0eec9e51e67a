"""The batched DSFW pieces against the plain host paths they replace: the one-launch
orientation choice, the one-launch backward elimination and the split-precision SYRK in
the streamed normal equations."""

import sys
import pytest
import torch
from test_solver_gpu import aperture_grid, close, fmcw, fmcw_specs, surfels  # noqa: F401

pytestmark = pytest.mark.gpu


def say(*args):
    print(*args, file=sys.stderr, flush=True)


@pytest.mark.parametrize(
    "mode,use_table,fresnel_n",
    [("complex", False, 0.0), ("real", False, 0.0), ("complex", True, 0.0), ("complex", True, 1.6),
     ("real", False, 1.6)],
)  # fmt: skip
def test_batched_orientation_choice_equals_the_chunked_loop(fmcw, mode, use_table, fresnel_n):  # noqa: F811
    from dataclasses import replace

    from dsplat.certificate import fibonacci_sphere, orient_candidates

    model, params, _ = fmcw
    model = replace(model, fresnel_n=fresnel_n)
    g = torch.Generator().manual_seed(11)
    count = 3000
    candidates = dict(
        centers=(torch.tensor([0, 0, 0.3]) + 0.02 * (2 * torch.rand(count, 3, generator=g) - 1)).cuda(),
        normals=torch.nn.functional.normalize(torch.randn(count, 3, generator=g), dim=1).cuda(),
        areas=torch.full((count,), 1e-4).cuda() * torch.linspace(0.5, 2, count).cuda(),
    )
    residual = model.scan.data
    table = model.range_table(residual, 8) if use_table else None
    plain, plain_scores = orient_candidates(model, candidates, residual, 24, mode, table, chunk=700)
    fused, fused_scores = orient_candidates(model, candidates, residual, 24, mode, table, batched=True)
    # the fused pass sums apertures in float32 in order, the chunked one per 32-aperture
    # block in float64: scores agree to float32 accuracy, and the fused choice scores within
    # that of the best under the chunked sums
    torch.testing.assert_close(fused_scores, plain_scores, rtol=2e-5, atol=1e-30)
    bank = fibonacci_sphere(24, "cuda")
    corr, norm = model.correlate_orientations(candidates, residual, bank, table)
    s = corr.real.double() if mode == "real" else corr.to(torch.complex128)
    values = s.abs().square() / norm.double().clamp(min=1e-60)
    normals = torch.cat((candidates["normals"][:, None], bank[None].expand(count, -1, -1)), 1)
    chosen = (normals == fused["normals"][:, None]).all(-1).float().argmax(1)
    picked = values.gather(1, chosen[:, None])[:, 0]
    assert torch.all(picked >= values.max(1).values * (1 - 2e-5))
    assert (chosen == values.argmax(1)).float().mean() > 0.99


@pytest.mark.parametrize(
    "n,count,limit", [(300, 40, None), (300, 40, 5), (2304, 256, None), (37, 37, None), (2500, 100, 0)]
)
def test_elimination_kernel_equals_the_host_loop(n, count, limit):
    from dsplat.replacement import backward_elimination

    g = torch.Generator().manual_seed(n + count)
    a = torch.randn(n + 50, n, generator=g, dtype=torch.float64)
    h = a.T @ a / n + 1e-3 * torch.eye(n, dtype=torch.float64)
    inverse = torch.linalg.inv(h).cuda()
    rhs = torch.randn(n, generator=g, dtype=torch.float64).cuda()
    limited = None if limit is None else torch.arange(n, device="cuda") < n // 2
    host = backward_elimination(inverse, rhs, count, limited, limit or 0)
    kernel = backward_elimination(inverse, rhs, count, limited, limit or 0, kernel=True)
    assert torch.equal(host, kernel)
    assert len(set(kernel.tolist())) == count
    if limit is not None:
        assert int((kernel < n // 2).sum()) <= limit


@pytest.mark.parametrize("gram", ["bf16x3", "fp16x3"])
def test_split_syrk_normal_equations_match_the_fp32_stream(fmcw, gram):  # noqa: F811
    from dsplat import DSFWConfig
    from dsplat.varpro import normal_equations

    model, params, _ = fmcw
    model = model.prepare(fmcw_specs())
    config = DSFWConfig(row_chunk=2 * model.shape[1])
    plain = normal_equations(model, params, model.scan.data, config)
    split = normal_equations(model, params, model.scan.data, DSFWConfig(row_chunk=2 * model.shape[1], gram=gram))
    close(split[0], plain[0], 1e-4 if gram == "bf16x3" else 1e-5)
    assert torch.equal(split[1], plain[1])


def test_batched_normal_equations_take_the_rhs_from_one_correlation(fmcw):  # noqa: F811
    from dsplat import DSFWConfig
    from dsplat.varpro import normal_equations

    model, params, _ = fmcw
    model = model.prepare(fmcw_specs())
    plain = normal_equations(model, params, model.scan.data, DSFWConfig(row_chunk=2 * model.shape[1]))
    fused = normal_equations(
        model, params, model.scan.data, DSFWConfig(row_chunk=2 * model.shape[1], gram="fp16x3", batched=True)
    )
    close(fused[0], plain[0], 1e-5)
    close(fused[1], plain[1], 1e-5)


@pytest.mark.parametrize("lockstep", [False, True])
@pytest.mark.parametrize("mode,normals", [("real", True), ("complex", True), ("complex", False)])
def test_qn_kernel_follows_the_host_ascent(fmcw, mode, normals, lockstep, monkeypatch):  # noqa: F811
    """Same candidates, same settings: the device ascent (one launch, or lockstep rounds on
    the sample-domain kernel) and the host loop end at the same certificates (their sums are
    rounded in different orders, so trajectories agree to float32 accuracy rather than bit
    for bit), and both reach the true reflector."""
    from dsplat import certificate
    from dsplat.certificate import CertificateSearchConfig, search_certificate

    model, params, coefficients = fmcw
    monkeypatch.setattr(certificate, "QN_MIN_BATCH", 1)
    monkeypatch.setattr(type(model), "lockstep_search", property(lambda self: lockstep))
    one = {k: v[:1] for k, v in params.items()}
    residual = model.render(one, coefficients[:1])
    start = {k: v.repeat(24, *([1] * (v.ndim - 1))) for k, v in one.items()}
    g = torch.Generator("cuda").manual_seed(1)
    start["centers"] = start["centers"] + 6e-4 * torch.randn(start["centers"].shape, device="cuda", generator=g)
    start["normals"] = torch.nn.functional.normalize(
        start["normals"] + 0.2 * torch.randn(start["normals"].shape, device="cuda", generator=g), dim=1
    )
    specs = fmcw_specs()
    config = CertificateSearchConfig(peaks=24, separation=0, iterations=25)
    kwargs = dict(normal_spec=specs["normals"] if normals else None)
    host, host_scores = search_certificate(model, residual, start, specs["centers"], config, mode, **kwargs)
    fused, fused_scores = search_certificate(
        model, residual, start, specs["centers"], config, mode, batched=True, **kwargs
    )
    torch.testing.assert_close(fused_scores, host_scores, rtol=2e-3, atol=1e-6)
    assert float((fused["centers"][0] - one["centers"][0]).norm()) < 5e-5
    assert float((host["centers"][0] - one["centers"][0]).norm()) < 5e-5
    if normals:
        torch.testing.assert_close(fused["normals"].norm(dim=1), torch.ones(len(fused["normals"]), device="cuda"))


class _OneKernel:
    """The model's certificate sums (with area), all from the sample-domain certificate
    gradient kernel, so that the host ascent compares values of the kernel the lockstep
    rounds evaluate."""

    def __init__(self, model, residual):
        from dsplat.kernels import surfel

        self.positions = model.sensor[1]
        self.operands = surfel.CertificateOperands(model.sensor[0], residual.reshape(model.shape), model.fresnel_n)

    def _sums(self, params):
        from dsplat.kernels import surfel

        out = surfel.certificate_gradient(self.positions, params["centers"], params["normals"], self.operands)
        area = params["areas"].float()
        scale = area.sqrt()
        corr = torch.complex(out[:, 0] * scale, out[:, 1] * scale)
        d_corr = torch.complex(out[:, 3:9] * scale[:, None], out[:, 9:15] * scale[:, None])
        return corr, out[:, 2] * area, d_corr, out[:, 15:21] * area[:, None]

    def correlate(self, params, residual, scan=True):
        return self._sums(params)[:2]

    def correlate_gradient(self, params, residual, include_normals=False):
        corr, norm, d_corr, d_norm = self._sums(params)
        return (corr, norm, d_corr, d_norm) if include_normals else (corr, norm, d_corr[:, :3], d_norm[:, :3])


@pytest.mark.parametrize(
    "mode,normals,fresnel_n,iterations,backtracks",
    [("complex", True, 0.0, 1, 8), ("complex", True, 0.0, 3, 8), ("real", True, 1.6, 3, 8),
     ("complex", False, 0.0, 3, 8), ("complex", True, 0.0, 8, 8), ("complex", True, 0.0, 0, 8),
     ("complex", True, 1.6, 3, 1)],
)  # fmt: skip
def test_lockstep_ascent_makes_the_host_loops_decisions(fmcw, mode, normals, fresnel_n, iterations, backtracks):  # noqa: F811
    """The lockstep rounds against the host loop scoring on the same evaluation kernel: every
    acceptance and halving compares values of one kernel, and the first iteration ends at
    the same points bit for bit. Later iterations differ by the rounding of the BFGS update
    (a batched matmul on the host, sequential sums here), which the ascent carries forward:
    about 1e-7 after two iterations, a few 1e-6 after eight. Also the stopping rules: no
    iterations, and a single trial per iteration."""
    from dataclasses import replace

    from dsplat.certificate import CertificateSearchConfig, _ascend_on_host, _SearchSpace, certificate
    from dsplat.kernels import search

    model, params, coefficients = fmcw
    model = replace(model, fresnel_n=fresnel_n)
    residual = model.render(params, coefficients) * 0.5 + model.render({k: v[:1] for k, v in params.items()},
                                                                         coefficients[:1])  # fmt: skip
    g = torch.Generator("cuda").manual_seed(3)
    pick = torch.arange(40, device="cuda") % len(params["centers"])
    start = {k: v[pick].clone() for k, v in params.items()}
    start["centers"] = start["centers"] + 8e-4 * torch.randn(start["centers"].shape, device="cuda", generator=g)
    start["normals"] = torch.nn.functional.normalize(
        start["normals"] + 0.3 * torch.randn(start["normals"].shape, device="cuda", generator=g), dim=1
    )
    specs = fmcw_specs()
    config = CertificateSearchConfig(peaks=40, separation=0, iterations=iterations, backtracks=backtracks)
    space = _SearchSpace(start, specs["centers"], specs["normals"] if normals else None)
    one = _OneKernel(model, residual)

    def score(p, gradient=False, scan=True):
        return certificate(one, p, residual, mode, gradient=gradient, normals=space.with_normals, scan=scan)

    host, host_values = _ascend_on_host(space, {k: v.clone() for k, v in space.start.items()}, score, config)
    c, n, values = search.qn_lockstep(model.sensor[1], space.start["centers"], space.start["normals"],
                                      space.start["areas"], space.kernel_spec(config, mode, residual), one.operands)
    begin = score(space.start)
    assert torch.all(values >= begin * (1 - 1e-6))
    if iterations == 0:
        assert torch.equal(c, space.start["centers"]) and torch.equal(n, space.start["normals"])
        torch.testing.assert_close(values, begin, rtol=1e-6, atol=0)
        return
    assert float((values / begin).max()) > 1.001  # the ascent moved
    if iterations == 1:
        assert torch.equal(values, host_values) and torch.equal(c, host["centers"])
        return
    rel = ((values - host_values) / host_values).abs()
    say(f"  lockstep vs host after {iterations}: values {float(rel.max()):.2e}, "
        f"centres {float((c - host['centers']).abs().max()):.2e}, normals {float((n - host['normals']).abs().max()):.2e}")
    if iterations <= 3:
        # centres near 0.3 m are float32 with a spacing of 3e-8: two ulps after three iterations
        torch.testing.assert_close(values, host_values, rtol=1e-6, atol=0)
        torch.testing.assert_close(c, host["centers"], rtol=0, atol=6e-8)
        torch.testing.assert_close(n, host["normals"], rtol=0, atol=2.4e-5)
        return
    # over eight iterations a candidate near an acceptance or halving decision can take the
    # other branch on that rounding: nearly all stay within a few 1e-6, and every one ends on
    # the same maximum to 1e-3
    same = rel <= 3e-5
    assert int(same.sum()) >= len(values) - 2, rel
    assert float(rel.max()) <= 1e-3
    torch.testing.assert_close(c[same], host["centers"][same], rtol=0, atol=2e-6)
    torch.testing.assert_close(n[same], host["normals"][same], rtol=0, atol=8e-4)
    if not normals:
        assert torch.equal(n, space.start["normals"])


@pytest.mark.parametrize("gram", ["plain", "fp16x3", "bf16x3"])
def test_batched_pooled_statistics_match_a_full_stream(fmcw, gram):  # noqa: F811
    from dsplat import DSFWConfig, initialize
    from dsplat.replacement import pooled_statistics
    from dsplat.varpro import normal_equations

    model, params, _ = fmcw
    config = DSFWConfig(gram=gram, batched=True)
    state = initialize(model, model.scan.data, surfels(5, seed=2), config, parameters=fmcw_specs())
    pool = {k: torch.cat((v, params[k])) for k, v in state.params.items()}
    gram_, rhs = pooled_statistics(state, pool, config)
    expected_gram, expected_rhs = normal_equations(state.model, pool, model.scan.data, DSFWConfig())
    close(gram_, expected_gram, 1e-4)
    close(rhs, expected_rhs, 1e-4)


@pytest.mark.parametrize("mode", ["fp32", "fp16x3"])
@pytest.mark.parametrize("m,n,col0", [(3000, 300, 256), (700, 130, 0), (5000, 513, 500), (64, 200, 199)])
def test_gemm_tn_is_the_trailing_cross_product(m, n, col0, mode):
    from dsplat.kernels.blas import gemm_tn

    g = torch.Generator().manual_seed(m + n + col0)
    rows = torch.randn(m, n, generator=g)
    out = torch.randn(n, n - col0, generator=g)
    got = gemm_tn(rows.cuda(), col0, out.clone().cuda(), mode).cpu().double()
    want = out.double() + rows.double().T @ rows.double()[:, col0:]
    assert ((got - want).abs().max() / want.abs().max()) < 2e-6


@pytest.mark.parametrize("pad,fresnel_n", [(2, 0.0), (2, 1.6), (1, 0.0), (4, 0.0)])
def test_projected_design_equals_the_projected_dense_design(pad, fresnel_n):
    """The fused kernel's rows equal the interleaved basis times the real bin-major design
    (design kernel, then reduce_major), to FP16 hi + lo accuracy."""
    from dsplat import FMCWConfig
    from dsplat.fmcw import response_basis
    from dsplat.kernels import surfel

    cfg = FMCWConfig(pad_factor=pad)
    positions = aperture_grid(6)
    p = surfels(300, seed=pad, spread=0.03)
    p["areas"] = p["areas"] * torch.linspace(0.5, 2.0, 300, device="cuda")
    k0 = (torch.cdist(p["centers"], positions) / cfg.range_bin_spacing_m)
    basis = response_basis(cfg, float(k0.min()) - 1, float(k0.max()) + 1, "cuda")
    assert basis is not None
    got = surfel.project_design(cfg, positions, p["centers"], p["normals"], p["areas"], basis.interleaved, fresnel_n)
    dense = surfel.design(cfg, positions, p["centers"], p["normals"], p["areas"], fresnel_n, True)
    want = basis.reduce_major(dense)
    assert got.shape == want.shape
    err = (got.double() - want.double()).norm() / want.double().norm()
    assert err < 2e-6, err


@pytest.mark.parametrize(
    "samples,pad,fresnel_n,count,spread",
    [(128, 2, 0.0, 300, 0.03), (128, 2, 1.6, 515, 0.15), (256, 2, 0.0, 260, 0.03), (128, 4, 0.0, 257, 0.05),
     (256, 1, 1.6, 100, 0.1), (128, 1, 0.0, 33, 0.01)],
)  # fmt: skip
def test_sampled_projection_matches_the_float64_projected_design(samples, pad, fresnel_n, count, spread):
    """Sample-domain projection (B = U^H F on the powers v^n) against the float64 oracle's real
    design times the same float32 basis; its sample basis against B formed from the complex
    U^H and e^{-2 pi i n k / N} directly. 25 apertures leave a last unit of one aperture; the
    surfel counts leave partial 256-column tiles; the wide spreads give ranks above 64 (a
    second 128-row pass)."""
    import numpy as np
    from dsplat import FMCWConfig
    from dsplat.fmcw import response_basis
    from dsplat.kernels import surfel
    from oracle import surfel as O

    cfg = FMCWConfig(adc_samples=samples, pad_factor=pad)
    positions = aperture_grid(5)
    p = surfels(count, seed=samples + pad + count, spread=spread)
    p["areas"] = p["areas"] * torch.linspace(0.5, 2.0, count, device="cuda")
    k0 = torch.cdist(p["centers"], positions) / cfg.range_bin_spacing_m
    basis = response_basis(cfg, float(k0.min()) - 1, float(k0.max()) + 1, "cuda")
    assert basis is not None
    rows, K, N = basis.interleaved.shape[0], cfg.num_range_bins, cfg.n_fft
    sampled = surfel.sample_basis(cfg, basis.interleaved, fresnel_n)
    assert sampled is not None and sampled.shape == (rows, 2 * samples)

    b = basis.interleaved.double().cpu().numpy()
    uh = b[0::2, 0::2] - 1j * b[0::2, 1::2]  # row 2j holds (Re, -Im) of U^H
    f = np.exp(-2j * np.pi * np.outer(np.arange(K), np.arange(samples)) / N)
    bs = uh @ f
    want_b = np.empty((rows, 2 * samples))
    want_b[0::2, 0::2], want_b[0::2, 1::2] = bs.real, -bs.imag
    want_b[1::2, 0::2], want_b[1::2, 1::2] = bs.imag, bs.real
    got_b = sampled.double().cpu().numpy()
    assert np.linalg.norm(got_b - want_b) / np.linalg.norm(want_b) < 1e-7

    maxima = torch.zeros(count, dtype=torch.int32, device="cuda")
    projected = surfel.project_design(
        cfg, positions, p["centers"], p["normals"], p["areas"], basis.interleaved, fresnel_n, sampled, maxima
    )
    # the column maxima formed on the way are exactly those of the written block, and the
    # FP16 SYRK given them equals the one that finds them itself
    assert torch.equal(maxima.view(torch.float32), projected.abs().amax(0))
    from dsplat.kernels.blas import syrk

    own = syrk(projected, torch.zeros(count, count, device="cuda"), "fp16x3")
    given = syrk(projected, torch.zeros(count, count, device="cuda"), "fp16x3", maxima)
    assert torch.equal(own, given)
    got = projected.double().cpu().numpy()
    dense = O.design(
        positions.cpu().numpy(), p["centers"].cpu().numpy(), p["normals"].cpu().numpy(), p["areas"].cpu().numpy(),
        cfg, fresnel_n, True,
    )  # fmt: skip
    want = (b @ dense.reshape(2 * K, -1)).reshape(rows * len(positions), count)
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    assert err < 2e-6, (rows, err)
    binned = surfel.project_design(cfg, positions, p["centers"], p["normals"], p["areas"], basis.interleaved, fresnel_n)
    assert np.linalg.norm(binned.double().cpu().numpy() - got) / np.linalg.norm(want) < 3e-6


def test_fused_normal_equations_match_the_plain_stream(fmcw):  # noqa: F811
    from dsplat import DSFWConfig
    from dsplat.varpro import normal_equations

    model, params, _ = fmcw
    model = model.prepare(fmcw_specs())
    plain = normal_equations(model, params, model.scan.data, DSFWConfig(row_chunk=2 * model.shape[1]))
    # bf16x3 in the batched form streams the projection kernel's split BF16 blocks into the
    # split SYRK (the chirp has 128 samples, so the sample-domain projection serves it)
    assert model.split_blocks(params, 3 * model.shape[1]) is not None
    for gram in ("plain", "fp16x3", "bf16x3"):
        fused = normal_equations(
            model, params, model.scan.data, DSFWConfig(row_chunk=3 * model.shape[1], gram=gram, batched=True)
        )
        close(fused[0], plain[0], 1e-5)
        close(fused[1], plain[1], 1e-5)


def test_pooled_cross_products_of_the_split_blocks_match_the_streamed_ones(fmcw):  # noqa: F811
    """The cross products of new columns with all columns (the pooled statistics) from the
    split BF16 blocks against float64 products of the float32 projected design."""
    from dsplat import DSFWConfig
    from dsplat.varpro import _Folded, block_products

    model, params, _ = fmcw
    model = model.prepare(fmcw_specs())
    count = len(params["centers"])
    col0 = count // 2
    acc = _Folded(count, count - col0, "cuda")
    config = DSFWConfig(row_chunk=3 * model.shape[1], gram="bf16x3", batched=True)
    assert block_products(model, params, model.scan.data, config, acc, col0=col0) is None
    dense = torch.cat([a for a in model.projected_blocks(params, 3 * model.shape[1])]).double()
    close(acc.result(), dense.T @ dense[:, col0:], 1e-5)

