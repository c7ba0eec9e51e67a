"""Both splat forms (sample-domain MFMA kernels and bin-domain cotangent kernels) against the
float64 oracle, for the chirps the sample-domain form covers (128 and 256 ADC samples, every
pad factor) and for ragged sizes: aperture and surfel counts that are not multiples of the
kernels' tiles, a single pair, back-facing surfels.

Bar: 1e-5 relative L2 against float64 for render and VJP, as the other splat tests.
"""

import sys

import numpy as np
import pytest
import torch
from conftest import aperture_grid, rel_l2, surfels

from oracle import surfel as O

pytestmark = pytest.mark.gpu
TOL = 1e-5


def say(*args):
    print(*args, file=sys.stderr, flush=True)


@pytest.fixture(params=["samples", "bins"])
def form(request):
    """The splat form each call of the test passes in its chirp (FMCWConfig.splat_form)."""
    return request.param


def tangents(count, seed):
    g = np.random.default_rng(seed + 200)
    dc = (1e-4 * g.standard_normal((count, 3))).astype(np.float32)
    dn = (1e-2 * g.standard_normal((count, 3))).astype(np.float32)
    dw = ((g.standard_normal(count) + 1j * g.standard_normal(count)) * 1e-3).astype(np.complex64)
    return dc, dn, dw


def run(cfg, positions, centers, normals, weights, upstream, fresnel_n, tangent=None):
    from dsplat.kernels import surfel

    d = torch.device("cuda", 0)
    p, c, n = (torch.as_tensor(x, dtype=torch.float32).to(d) for x in (positions, centers, normals))
    w = torch.as_tensor(weights, dtype=torch.complex64).to(d)
    u = torch.as_tensor(upstream, dtype=torch.complex64).to(d)
    out = surfel.render(cfg, p, c, n, w, fresnel_n)
    gc, gn, gw = surfel.vjp(cfg, p, c, n, w, u, fresnel_n)
    result = [out, gc, gn, gw]
    if tangent is not None:
        dc, dn, dw = (torch.as_tensor(x).to(d) for x in tangent)
        result.append(surfel.jvp(cfg, p, c, n, w, dc, dn, dw, fresnel_n))
    torch.cuda.synchronize()
    return [x.cpu().numpy() for x in result]


def scene(count, apertures_side, seed, back_facing=0):
    s = surfels(count, seed=seed, spread=0.012)
    normals = s["normals"].copy()
    if back_facing:
        normals[:back_facing] *= -1.0  # these see cos <= 0 from every aperture
    g = np.random.default_rng(seed)
    weights = (g.standard_normal(count) + 1j * g.standard_normal(count)).astype(np.complex64) * 1e-2
    positions = aperture_grid(apertures_side, 0.003)
    return positions, s["centers"], normals, weights


def check(cfg, positions, centers, normals, weights, fresnel_n, seed):
    g = np.random.default_rng(seed + 100)
    up = (g.standard_normal((len(positions), cfg.num_range_bins))
          + 1j * g.standard_normal((len(positions), cfg.num_range_bins))).astype(np.complex64)  # fmt: skip
    tangent = tangents(len(centers), seed)
    out, gc, gn, gw, dy = run(cfg, positions, centers, normals, weights, up, fresnel_n, tangent)
    args = (positions, centers, normals, weights)
    want = O.render(*args, cfg, fresnel_n)
    ec, en, ew = O.vjp(*args, up, cfg, fresnel_n)
    edy = O.jvp(*args, *tangent, cfg, fresnel_n)
    errs = dict(render=rel_l2(out, want), gc=rel_l2(gc, ec), gn=rel_l2(gn, en), gw=rel_l2(gw, ew),
                jvp=rel_l2(dy, edy))  # fmt: skip
    say(f"samples={cfg.adc_samples} pad={cfg.pad_factor} A={len(positions)} S={len(centers)} "
        f"n={fresnel_n}: " + " ".join(f"{k} {v:.2e}" for k, v in errs.items()))  # fmt: skip
    assert max(errs.values()) < TOL, errs


@pytest.mark.parametrize(
    "samples,pad,fresnel_n",
    [(128, 2, 0.0), (128, 1, 1.6), (128, 4, 0.0), (128, 8, 1.6), (256, 2, 0.0), (256, 1, 0.0), (256, 4, 1.6)],
)
def test_forms_match_the_oracle_across_chirps(form, samples, pad, fresnel_n):
    from dsplat import FMCWConfig

    cfg = FMCWConfig(adc_samples=samples, pad_factor=pad, splat_form=form)
    check(cfg, *scene(40, 9, seed=samples + pad), fresnel_n, seed=pad)


@pytest.mark.parametrize("side,count,back", [(1, 1, 0), (5, 17, 3), (7, 33, 0), (4, 2, 1)])
def test_forms_on_ragged_sizes_and_back_faces(form, side, count, back):
    from dsplat import FMCWConfig

    cfg = FMCWConfig(splat_form=form)
    check(cfg, *scene(count, side, seed=side * 31 + count, back_facing=back), 0.0, seed=count)


def test_the_two_forms_agree_on_a_larger_case():
    from dsplat import FMCWConfig

    args = scene(300, 24, seed=7)
    g = np.random.default_rng(3)
    up = (g.standard_normal((24 * 24, 128)) + 1j * g.standard_normal((24 * 24, 128))).astype(np.complex64)
    results = [run(FMCWConfig(splat_form=f), *args, up, 0.0, tangents(300, 7)) for f in ("samples", "bins")]
    errs = [rel_l2(a, b) for a, b in zip(results[0], results[1])]
    say("sample-domain vs bin-domain rel L2: " + " ".join(f"{e:.2e}" for e in errs))
    assert max(errs) < 2 * TOL
