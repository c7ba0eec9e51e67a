"""The split-operand normal equations: the sample-domain projection written as 16-bit hi / lo
blocks in 32-row column tiles, and the SYRK that multiplies such blocks (hi hi + hi lo + lo hi
on the BF16 or FP16 matrix cores), against float64 products and against the float32
projection.

Bars: hi + lo within 2^-17 (BF16) or 2^-21 plus half an FP16 subnormal step (FP16, columns
scaled by powers of two) of the float32 value per element; products within 1e-5 (BF16) or
1e-6 (FP16) relative Frobenius of float64 (the omitted lo lo term and the lo rounding, FP32
accumulation over m rows); untouched lower triangle; ragged sizes (n not a multiple of the
256-column tile, one 32-row stage, a diagonal-only tile).
"""

import sys

import pytest
import torch

pytestmark = pytest.mark.gpu


def say(*args):
    print(*args, file=sys.stderr, flush=True)


@pytest.mark.parametrize("fmt", ["bf16", "fp16"])
@pytest.mark.parametrize("m,n,col0", [(32, 5, None), (4096, 300, None), (65536, 2035, None), (8192, 700, 600),
                                      (2048, 513, 0), (96, 257, 256)])
def test_syrk_split_matches_float64(m, n, col0, fmt):
    from dsplat.kernels import blas

    g = torch.Generator().manual_seed(m + n)
    scales = torch.exp(2 * torch.randn(n, generator=g))  # columns over a wide range of magnitudes
    rows = (torch.randn(m, n, generator=g) * scales * 1e-6).cuda()
    # FP16: each column scaled by a power of two that puts its largest entry just below 2^14
    exponents = torch.floor(torch.log2(2.0**14 / rows.abs().amax(0))).float() if fmt == "fp16" else None
    block = blas.SplitBlock.of(rows, fmt, exponents)
    tail = n if col0 is None else n - col0
    out = torch.zeros((n, tail), device="cuda")
    if col0 is None:
        out = torch.tril(torch.full((n, n), 3.0, device="cuda"), -1)  # a sentinel below the diagonal
    say(f"syrk_split {fmt} m {m} n {n} col0 {col0}")
    blas.syrk_split(block, out, col0)
    dense = block.dense().double()
    want = dense.T @ (dense if col0 is None else dense[:, col0:])
    got = out.double()
    if col0 is None:
        upper = torch.triu(torch.ones(n, n, dtype=torch.bool, device="cuda"))
        assert torch.all(out[~upper] == 3.0)  # the lower triangle is untouched
        diff, ref = (got - want)[upper], want[upper]
    else:
        diff, ref = got - want, want
    err = float(diff.norm() / ref.norm())
    say(f"  {fmt} rel err {err:.2e}")
    assert err < (1e-5 if fmt == "bf16" else 1e-6)


@pytest.mark.parametrize("fmt", ["bf16", "fp16"])
def test_split_projection_reconstructs_the_float32_projection(fmt):
    from dsplat import FMCWConfig
    from dsplat.groundtruth import planar_array
    from dsplat.kernels import blas, surfel

    for samples, rows in ((128, 64), (256, 64), (128, 50)):
        cfg = FMCWConfig(adc_samples=samples)
        g = torch.Generator().manual_seed(samples + rows)
        K = cfg.num_range_bins
        basis = torch.linalg.qr(torch.randn(2 * K, rows, generator=g, dtype=torch.float64))[0].T.float().cuda()
        pos = planar_array(24, 0.0025).cuda()[:555]
        S = 300
        centers = (torch.tensor([0.0, 0.0, 0.35]) + 0.02 * (2 * torch.rand(S, 3, generator=g) - 1)).cuda()
        normals = torch.nn.functional.normalize(torch.randn(S, 3, generator=g) - torch.tensor([0, 0, 2.0]), dim=1).cuda()
        areas = torch.full((S,), 2.5e-6).cuda()
        sampled = surfel.sample_basis(cfg, basis)
        say(f"split projection {fmt}: {samples} samples, {rows} rows, {len(pos)} apertures x {S} surfels")
        dense = surfel.project_design(cfg, pos, centers, normals, areas, basis, 0.0, sampled=sampled)
        exponents = blas.projection_exponents(cfg, pos, centers, areas) if fmt == "fp16" else None
        block = surfel.project_split(cfg, pos, centers, normals, areas, sampled, rows, 0.0, fmt, exponents)
        pad = (rows + 31) // 32 * 32
        assert block.m == pad * len(pos)
        recon = block.dense().reshape(len(pos), pad, S)
        assert torch.all(recon[:, rows:] == 0)
        want = dense.reshape(rows, len(pos), S).permute(1, 0, 2)
        err = (recon[:, :rows] - want).abs()
        if fmt == "bf16":
            assert torch.all(err <= 2.0**-17 * want.abs() + 1e-30)
        else:
            # the a-priori scales bound every entry: the largest stays below 2^14 after scaling
            assert float((want.abs() * torch.exp2(exponents)).max()) < 2.0**14
            # two FP16 roundings: 2^-22 relative, and half the FP16 subnormal step once lo is subnormal
            assert torch.all(err <= 2.0**-21 * want.abs() + 2.0**-25 * torch.exp2(-exponents))
        # the split SYRK of the split block equals float64 products of the float32 block
        gram = torch.zeros(S, S, device="cuda")
        blas.syrk_split(block, gram)
        ref = dense.double().T @ dense.double()
        upper = torch.triu(torch.ones(S, S, dtype=torch.bool, device="cuda"))
        rel = float((gram.double() - ref)[upper].norm() / ref[upper].norm())
        say(f"  gram rel err {rel:.2e}")
        assert rel < (1e-5 if fmt == "bf16" else 1e-6)
