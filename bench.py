"""Benchmark of the Dirichlet-splat hot path on one MI355X.

The metric has two halves.

* Splat: a step is one forward render plus one VJP of the full-size scene: 128x128
  apertures at 2.5 mm, 128 range bins (120 GHz, 15 GHz, 128 ADC samples, pad 2), 2048
  front-facing surfels on the 30/30/25 mm ellipsoid at 0.35 m. `value` counts (aperture,
  bin, surfel) evaluations per second of fwd+VJP; inputs are resident in HBM before the
  timed region.
* DSFW: the full fit (float64 truth scan of the front faces of 4000 ellipsoid draws, 256 ADC
  samples, 30 dB noise, 100 iterations, batch 256) in the plain FP32 form and in the
  mixed-precision batched form, as the `dsfw` object: median step ms of each, the speed-up,
  the NMSE trajectories' deviation with its perturbed-target control, final NMSE and median
  centre error.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--ramp-ms T] [--dsfw-steps D] [--no-cpu]

Prints one JSON line: value, ms_per_step, a roofline object for the dominant splat launch
group (HIP events on the launch stream), the dsfw object, and as cpu_baseline the
reference's float32 model in torch on every host core (the float64 oracle beside it).
"""

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "dirichlet-splatting_amd"))

# Algorithmic FP32 work per (a, k, s) evaluation.
# Bin-domain (cotangent) kernels, VALU:
#   forward 13 FLOP: cot by angle addition (3 + 3), cot (1), complex accumulate (4 + 2)
#   VJP     19 FLOP: the forward's 13 plus the (1 + cot^2) u class sum (6)
#   scan    11 FLOP: cot by angle addition (5), residual accumulate (4), |cot|^2 (2)
# Sample-domain kernels, MFMA (128 samples = 128 bins at pad 2, so per (a, n, s) = per eval).
# The samples of a row j are split symmetrically (n = 16 j + 7 +- i, or 7.5 +- (i + 1/2)), so
# conjugate pairs of powers share their products: the counts are the real multiply-adds of
# these sums, FP32-accurate (FP16 hi + lo operands, three products, FP32 accumulation) and so
# priced against the FP32 peak.
#   forward  4.5 FLOP: per (a, s, row of 16 samples) 9 W columns x 4 real products
#   VJP      8 FLOP: per (a, s, row) 8 W_i x (Re, Im) x 2 output parts x 2 sums (w^n U, n w^n U),
#            the symmetric-split kernel of pad 1; at pad >= 2 the VJP runs by bin moments, with no
#            per-(a, n, s) sums, and is measured by its VALU issue
#   scan     8 FLOP: as the VJP, the residual's sum and the norm polynomial's
FWD_FLOP, VJP_FLOP, SCAN_FLOP = 13, 19, 11
TD_FWD_FLOP, TD_VJP_FLOP = 4.5, 8
TD_SCAN_FLOP = 8
# FP16 MFMA work the sample-domain kernels execute per (a, s) pair: six 16x16x32 (forward) or
# three 32x32x16 (VJP) products per aperture and 32 surfels
TD_EXEC_FLOP_PER_PAIR = 3072
PEAK_FP16_TFLOPS = 2516.6  # MI355X published dense FP16 MFMA peak
PEAK_FP32_TFLOPS = 157.3  # MI355X published FP32 peak (vector and dense MFMA)
PEAK_HBM_GBS = 8000.0
TRAFFIC = ROOT / "profiles" / "splat_traffic_128x128x2048.json"
# VALU instructions per launch of the headline scene (rocprofv3 SQ_INSTS_VALU per dispatch)
VALU_PROFILE = ROOT / "profiles" / "pmc_valu_splat_128x128x2048.md"
VALU_PER_LAUNCH = {"td_render_h16_kernel<false, false>": 183_625_072, "vjp_moments_kernel<false>": 112_459_776}
# the published FP32 peak is one wave64 FP32 VALU instruction per 2 cycles on each SIMD
VALU_PEAK_PER_S = PEAK_FP32_TFLOPS * 1e12 / (64 * 2)
SIMDS, CLOCK_GHZ = 1024, 2.4  # MI355X: 256 CUs x 4 SIMDs, peak engine clock


def measured_traffic(kernels):
    """HBM bytes per launch of a kernel group from the committed rocprofv3 --pmc record
    (FETCH_SIZE and WRITE_SIZE in KiB per dispatch, scaled by what the same counters read
    for a plain copy of known size), or None if the record does not cover the group."""
    if not TRAFFIC.exists():
        return None
    record = json.loads(TRAFFIC.read_text())
    try:
        per = [record["kernels"][k] for k in kernels]
    except KeyError:
        return None
    fetch = sum(p["fetch_kib"] for p in per) * record["fetch_scale"]
    write = sum(p["write_kib"] for p in per) * record["write_scale"]
    return 1024.0 * (fetch + write)


def parse():
    parser = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=100)
    parser.add_argument("--warmup", type=int, default=3)
    parser.add_argument("--ramp-ms", type=float, default=250.0,
                        help="untimed steps for this much wall time before the warm-up (clocks settle)")
    parser.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    parser.add_argument("--cpu-seconds", type=float, default=8.0, help="target float64-oracle sample time")
    parser.add_argument(
        "--workload", choices=("splat", "gram", "scan", "dsfw", "synth", "syrk", "copy", "valu"), default="splat",
        help="splat: the fwd+VJP metric (default); gram: the split-precision Gram of the same scene; "
        "scan: the batched certificate scan of 64^3 candidates against 64^2, 96^2 and 128^2 scans; "
        "dsfw: DSFW steps, plain FP32 form against the mixed-precision batched form; "
        "synth: the float64 truth scan of the DSFW fit by the ADC signal chain on the FP64 matrix cores; "
        "syrk: the split-operand SYRK of the mixed DSFW normal equations at two row counts; "
        "copy: a plain 1 GiB device copy, the reference size for the HBM byte counters; "
        "valu: the FP32 and FP64 VALU FMA issue rate that bounds the splat kernels",
    )  # fmt: skip
    parser.add_argument("--adc-samples", type=int, default=256, help="dsfw: ADC samples (bins = samples)")
    parser.add_argument("--truth-draws", type=int, default=4000,
                        help="dsfw: area-uniform ellipsoid draws whose front faces are the truth")
    parser.add_argument("--front-count", type=int, default=0,
                        help="dsfw: exactly this many front faces as the truth instead (0: use --truth-draws)")
    parser.add_argument("--dsfw-form", choices=("both", "plain_fp32", "mixed_batched"), default="both",
                        help="--workload dsfw: compare both forms, or time one form's steps alone")
    parser.add_argument("--dsfw-steps", type=int, default=100,
                        help="default run: timed DSFW steps per form (0 skips the DSFW comparison)")
    return parser.parse_args()


def cpu_baseline(work, target_seconds):
    """The float64 numpy oracle (fwd + VJP) on a bounded aperture sample, one thread."""
    import numpy as np
    from threadpoolctl import threadpool_limits

    from oracle import surfel as O

    pos, c, n = work["positions"].numpy(), work["centers"].numpy(), work["normals"].numpy()
    w, up, cfg = work["weights"].numpy(), work["cotangent"].numpy(), work["config"]
    bins, count = cfg.num_range_bins, len(c)

    def run(ids):
        t0 = time.perf_counter()
        O.render(pos[ids], c, n, w, cfg)
        O.vjp(pos[ids], c, n, w, up[ids], cfg)
        return time.perf_counter() - t0

    with threadpool_limits(limits=1):
        probe = np.arange(0, len(pos), len(pos) // 8)[:8]
        t_probe = run(probe)
        apertures = int(min(256, max(8, 8 * target_seconds / max(t_probe, 1e-3))))
        ids = np.linspace(0, len(pos) - 1, apertures).astype(np.int64)
        seconds = run(ids)
    evals = len(ids) * bins * count
    return {
        "value": evals / seconds,
        "unit": "evals/s",
        "cores": 1,
        "kind": "port",
        "sample": f"float64 numpy oracle, render+VJP of {len(ids)} of {len(pos)} apertures x {bins} bins x "
        f"{count} surfels ({evals:.3e} evals, {seconds:.1f} s, 1 thread)",
    }


def torch_fp32_splat(positions, centers, normals, weights, cfg):
    """The reference's float32 direct-form splat as torch tensor operations, differentiable by
    autograd: the column model of surfel.slang:63-142 (amp = cos^2 lambda / (8 pi r) for cos > 0,
    the carrier phase 2 pi (fc tau + slope tau (t0 - tau / 2)) in float32) times the closed-form
    Dirichlet kernel of common.slang:10-45, D(x) = sin(pi S x / N) / (S sin(pi x / N))
    e^{-i pi (S - 1) x / N} with x = k - k0 wrapped to [-N/2, N/2) and its Taylor form near 0."""
    import math

    import torch

    S, N, K = cfg.adc_samples, cfg.n_fft, cfg.num_range_bins
    offset = positions[:, None, :] - centers[None]
    r = offset.norm(dim=-1)
    cosine = ((offset / r[..., None]) * normals[None]).sum(-1)
    amp = torch.where(cosine > 0, cosine.square() * (cfg.lambda_m / (8 * math.pi)) / r, torch.zeros_like(r))
    tau = 2 * r / cfg.c0_m_s
    phase = 2 * math.pi * (cfg.fc_hz * tau + cfg.slope_hz_s * tau * (cfg.adc_start_time_s - 0.5 * tau))
    coef = torch.polar(amp, phase) * weights[None]
    x = torch.arange(K, dtype=r.dtype)[None, None] - (r / cfg.range_bin_spacing_m)[..., None]
    x = torch.remainder(x + N / 2, N) - N / 2
    small = x.abs() < 1e-3
    den = torch.where(small, torch.ones_like(x), S * torch.sin(math.pi * x / N))
    ratio = torch.where(small, torch.ones_like(x), torch.sin(math.pi * S * x / N) / den)
    D = torch.polar(ratio.abs(), -math.pi * (S - 1) * x / N) * torch.sign(ratio)
    return (coef[..., None] * D).sum(1)


def cpu_torch_fp32(work, repeats=5):
    """The reference's float32 model in torch on every host core, render + autograd VJP, as the
    reported CPU baseline: the small 32x32-aperture, 200-surfel case in full and the first 1024
    apertures of the full-size scene (all 2048 surfels, 128 bins), one warm-up then `repeats`
    timed runs each, median wall time; plus the relative L2 error of its forward against the
    float64 oracle on 32 of those apertures."""
    import numpy as np
    import torch

    from dsplat import workloads
    from oracle import surfel as O

    torch.set_num_threads(os.cpu_count() or 1)
    cfg = work["config"]

    def run(pos, c, n, w, up):
        t0 = time.perf_counter()
        outs = []
        for chunk in torch.split(torch.arange(len(pos)), 8):
            cc = c.clone().requires_grad_(True)
            nn = n.clone().requires_grad_(True)
            ww = w.clone().requires_grad_(True)
            out = torch_fp32_splat(pos[chunk], cc, nn, ww, cfg)
            (up[chunk].conj() * out).real.sum().backward()
            outs.append(out.detach())
        return time.perf_counter() - t0, torch.cat(outs)

    def median_rate(pos, c, n, w, up):
        run(pos[:8], c, n, w, up[:8])  # first-call set-up
        times = [run(pos, c, n, w, up)[0] for _ in range(repeats)]
        return len(pos) * cfg.num_range_bins * len(c) / float(np.median(times)), times

    small = workloads.fmcw_small()
    small_rate, small_times = median_rate(*(small[k] for k in ("positions", "centers", "normals", "weights", "cotangent")))
    first = slice(0, 1024)
    pos, c, n, w, up = (work[k] for k in ("positions", "centers", "normals", "weights", "cotangent"))
    rate, times = median_rate(pos[first], c, n, w, up[first])
    _, out = run(pos[:1024:32], c, n, w, up[:1024:32])
    want = O.render(pos[:1024:32].numpy(), c.numpy(), n.numpy(), w.numpy(), cfg)
    rel = float(np.linalg.norm(out.numpy() - want) / np.linalg.norm(want))
    model = next((ln.split(":", 1)[1].strip() for ln in open("/proc/cpuinfo") if ln.startswith("model name")), "?")
    return {
        "value": rate, "unit": "evals/s", "cores": torch.get_num_threads(), "kind": "port",
        "sample": f"float32 torch restatement of the reference model (closed-form Dirichlet kernel, float32 carrier "
        f"phase), render + autograd VJP of the first 1024 of {len(pos)} apertures x {cfg.num_range_bins} bins x "
        f"{len(c)} surfels, median of {repeats} runs ({float(np.median(times)):.2f} s each)",
        "os_cpu_count": os.cpu_count(), "cpu": model, "torch": torch.__version__,
        "run_seconds": times, "rel_l2_forward_vs_float64": rel,
        "small_case": {"value": small_rate, "unit": "evals/s", "run_seconds": small_times,
                       "workload": "32x32 apertures @ 2.5 mm, 128 bins, 200 surfels in a 60 mm cube, in full"},
    }  # fmt: skip


def timed(fn, repeats, stream):
    """Mean milliseconds of fn() between HIP events on the launch stream, after one warm-up."""
    import torch

    fn()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record(stream)
    for _ in range(repeats):
        out = fn()
    stop.record(stream)
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / repeats, out


def gram_workload(args):
    """A^H A and A^H r of the full-size dictionary (16384 x 128 rows, 2048 surfel columns)
    in FP32, BF16 hi+lo and FP16 hi+lo MFMA, each against float64 products of the same
    float32 atoms, then an FP64 Cholesky solve and the coefficient error."""
    import torch

    from dsplat import workloads
    from dsplat.kernels import blas, surfel
    from dsplat.varpro import solve

    work = workloads.ellipsoid_scene()
    cfg = work["config"]
    dev = torch.device("cuda", 0)
    pos, c, n, areas = (work[k].to(dev) for k in ("positions", "centers", "normals", "areas"))
    A, K, S = len(pos), cfg.num_range_bins, len(c)
    M = A * K
    residual = workloads.complex_normal((M,), torch.Generator().manual_seed(3)).to(dev)
    stream = torch.cuda.current_stream(dev)
    # float64 reference: products of the float32 atoms accumulated in complex128
    g64 = torch.zeros((S, S), dtype=torch.complex128, device=dev)
    q64 = torch.zeros(S, dtype=torch.complex128, device=dev)
    chunk = 128
    for start in range(0, A, chunk):
        a = surfel.design(cfg, pos[start : start + chunk], c, n, areas, 0.0, False).to(torch.complex128)
        g64 += a.H @ a
        q64 += a.H @ residual[start * K : (start + chunk) * K].to(torch.complex128)
        del a
    g64, q64 = g64 / M, q64 / M
    eye = torch.eye(S, dtype=torch.complex128, device=dev)
    scale = float(g64.diagonal().real.mean())
    ridges = {"1e-4": 1e-4 * scale, "1e-2": 1e-2 * scale}  # relative to the mean diagonal
    b64 = {key: solve(g64 + ridge * eye, q64) for key, ridge in ridges.items()}

    def coefficient_errors(g, q):
        out = {}
        for key, ridge in ridges.items():
            b = solve(g + ridge * eye, q)
            d = b - b64[key]
            out[f"coef_rel_err_ridge{key}"] = float(d.norm() / b64[key].norm())
            # prediction error ||A (b - b64)|| / ||A b64||, well conditioned
            out[f"pred_rel_err_ridge{key}"] = float(
                ((d.conj() @ g64 @ d).real / (b64[key].conj() @ g64 @ b64[key]).real).sqrt()
            )
        return out
    q_ms, (corr, _) = timed(lambda: surfel.correlate(cfg, pos, c, n, residual.reshape(A, K), 0.0), 3, stream)
    q = (corr * areas.sqrt()).to(torch.complex128) / M
    useful = 4.0 * M * S * (S + 1)  # Hermitian half, complex MAC = 8 FLOP
    modes = {}
    for mode in ("fp32", "bf16x3", "fp16x3"):
        ms, g = timed(lambda mode=mode: blas.gram(cfg, pos, c, n, areas, mode, scale=1.0 / M), 3, stream)
        g = g.to(torch.complex128)
        executed = (1 if mode == "fp32" else 3) * 8.0 * M * (S * S / 2 + 64 * S)  # full upper tiles
        modes[mode] = {
            "ms": ms,
            "useful_tflops": useful / (ms * 1e-3) / 1e12,
            "mfma_tflops": executed / (ms * 1e-3) / 1e12,
            "gram_rel_err": float((g - g64).norm() / g64.norm()),
            **coefficient_errors(g, q),
        }
    peak = {"fp32": 157.3, "bf16x3": 2516.6, "fp16x3": 2516.6}
    best = "bf16x3"
    return {
        "metric": "VarPro Gram A^H A ms (16384x128 rows x 2048 surfels, complex)",
        "value": modes[best]["ms"],
        "unit": "ms",
        "n_gpus": 1,
        "higher_is_better": False,
        "dtype": "bf16 hi+lo split, fp32 accumulation",
        "data": "synthetic",
        "config": {"workload": "Gram of the 128x128 x 128-bin x 2048-surfel ellipsoid dictionary", "rows": M,
                   "columns": S, "ridges_relative_to_mean_diagonal": list(ridges)},
        "reference_fp64_coefficients_from_fp64_rhs": True,
        "rhs": {"ms": q_ms, "rel_err": float((q - q64).norm() / q64.norm()), "path": "correlate kernel, fp32 VALU"},
        "modes": modes,
        "speedup_bf16x3_vs_fp32": modes["fp32"]["ms"] / modes["bf16x3"]["ms"],
        "roofline": {"kernel": "gram_kernel<2,1>", "bound": "mfma", "achieved": modes[best]["mfma_tflops"],
                     "peak": peak[best], "unit": "TFLOP/s", "frac": modes[best]["mfma_tflops"] / peak[best],
                     "traffic": None},
    }


def torch_synthesis(positions, centers, normals, amplitudes, cfg, aperture_block=256, surfel_block=512):
    """The float64 beat-tone chain of groundtruth.py:58-90 as torch tensor operations over
    aperture and surfel blocks (the reference's own formulation): the comparison leg of the
    synthesis workload."""
    import math

    import torch

    t = torch.arange(cfg.adc_samples, device=centers.device, dtype=torch.float64) / cfg.sample_rate_hz
    centers, normals, amplitudes = centers.double(), normals.double(), amplitudes.to(torch.complex128)
    rows = []
    for a0 in range(0, len(positions), aperture_block):
        block = positions[a0 : a0 + aperture_block].double()
        adc = torch.zeros((len(block), 1, cfg.adc_samples), dtype=torch.complex128, device=centers.device)
        for s0 in range(0, len(centers), surfel_block):
            sl = slice(s0, s0 + surfel_block)
            offset = block[:, None, :] - centers[None, sl]
            r = offset.norm(dim=-1).clamp(min=1e-9)
            cosine = ((offset / r[..., None]) * normals[None, sl]).sum(-1).clamp(min=0)
            gain = cosine.square() * (cfg.lambda_m / (8 * math.pi)) / r
            delay = 2 * r / cfg.c0_m_s
            phase = 2 * math.pi * (cfg.fc_hz * delay + cfg.slope_hz_s * delay * (cfg.adc_start_time_s - 0.5 * delay))
            echo = torch.polar(gain, phase) * amplitudes[sl]
            beat = 2 * cfg.slope_hz_s * r / cfg.c0_m_s
            tones = torch.polar(torch.ones_like(r)[..., None], 2 * math.pi * beat[..., None] * t)
            adc += torch.bmm(echo[:, None, :], tones)
        rows.append(torch.fft.fft(adc[:, 0], n=cfg.n_fft)[:, : cfg.num_range_bins] / cfg.adc_samples)
    return torch.cat(rows)


def synth_workload(args):
    """The float64 truth scan of the DSFW fit: 128x128 apertures at 2.5 mm, 256 ADC samples
    (pad 2, 256 bins), the front faces of 4000 area-uniform ellipsoid draws with amplitude
    sqrt(area), made by the ADC signal chain on the FP64 matrix cores (dsplat_synthesize).
    Also times the reference's torch float64 block loop on the same inputs and checks the
    kernel against the float64 numpy oracle on a sample of apertures."""
    import numpy as np
    import torch
    from threadpoolctl import threadpool_limits

    from dsplat import FMCWConfig, workloads
    from dsplat.groundtruth import planar_array
    from dsplat.kernels import surfel

    dev = torch.device("cuda", 0)
    cfg = FMCWConfig(adc_samples=256)
    centers, normals = (x.to(dev) for x in workloads.ellipsoid_samples(4000, 0, keep_all_draws=True))
    area = workloads.ellipsoid_area(*workloads.SEMI_AXES) / 4000
    positions = planar_array(128, workloads.PITCH).to(dev)
    amplitudes = torch.full((len(centers),), area**0.5, dtype=torch.complex128, device=dev)
    A, S, T = len(positions), len(centers), cfg.adc_samples
    evals = A * T * S
    stream = torch.cuda.current_stream(dev)

    def step():
        return surfel.synthesize(cfg, positions, centers, normals, amplitudes, 0.0)

    for _ in range(max(1, args.warmup)):
        step()
    steps = max(1, min(args.steps, 20))
    kernel_ms, out = timed(step, steps, stream)
    sample = torch.arange(0, A, 1024, device=dev)  # 16 apertures across the array
    torch_ms, ref = timed(lambda: torch_synthesis(positions[sample], centers, normals, amplitudes, cfg), 1, stream)
    torch_ms_full = torch_ms * A / len(sample)
    from oracle import surfel as O

    want = O.synthesize(positions[sample].cpu().numpy(), centers.cpu().numpy(), normals.cpu().numpy(),
                        amplitudes.cpu().numpy(), cfg)  # fmt: skip
    got = out[sample].cpu().numpy()
    rel = float(np.linalg.norm(got - want) / np.linalg.norm(want))
    rel_torch = float(np.linalg.norm(ref.cpu().numpy() - want) / np.linalg.norm(want))
    achieved = 8 * evals / (kernel_ms * 1e-3) / 1e12
    result = {
        "metric": "FP64 scan synthesis evals/s (aperture·sample·surfel)",
        "value": evals / (kernel_ms * 1e-3),
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": max(1, args.warmup),
        "ms_per_step": kernel_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"float64 FMCW truth scan: 128x128 apertures @ 2.5 mm, 256 ADC samples (256 bins), "
                   f"{S} front-facing surfels of 4000 ellipsoid draws", "apertures": A, "samples": T,
                   "surfels": S, "evals_per_step": evals, "parallelism": "single GPU"},
        "kernels_ms": {"synthesize": kernel_ms, "torch_float64_blocks": torch_ms_full},
        "speedup_vs_torch_float64": torch_ms_full / kernel_ms,
        "rel_l2_vs_float64_oracle": {"kernel": rel, "torch_float64_blocks": rel_torch, "apertures": len(sample)},
        "roofline": {"kernel": "synthesize_kernel<16>", "bound": "mfma", "achieved": achieved, "peak": 78.6,
                     "unit": "TFLOP/s", "frac": achieved / 78.6, "traffic": None, "flop_per_eval": 8,
                     "peak_basis": "FP64 dense MFMA (one complex MAC per (aperture, sample, surfel))"},
    }
    if not args.no_cpu:
        pos, cc, cn = positions.cpu().numpy(), centers.cpu().numpy(), normals.cpu().numpy()
        amps = amplitudes.cpu().numpy()
        with threadpool_limits(limits=1):
            t = time.perf_counter()
            O.synthesize(pos[:2], cc, cn, amps, cfg)
            probe = time.perf_counter() - t
            count = int(min(A, max(2, 2 * args.cpu_seconds / max(probe, 1e-3))))
            ids = np.linspace(0, A - 1, count).astype(np.int64)
            t = time.perf_counter()
            O.synthesize(pos[ids], cc, cn, amps, cfg)
            seconds = time.perf_counter() - t
        result["cpu_baseline"] = {
            "value": count * T * S / seconds, "unit": "evals/s", "cores": 1, "kind": "port",
            "sample": f"float64 numpy oracle synthesis of {count} of {A} apertures ({count * T * S:.3e} evals, "
            f"{seconds:.1f} s, 1 thread)",
        }  # fmt: skip
    return result


def scan_workload(args):
    """Certificates |a_c^H r|^2 / (|a_c|^2 |r|^2) of 262144 candidates (64^3 grid over the
    70 mm box, seeded -z-facing normals) against three scans of uneven size (64^2, 96^2
    and 128^2 apertures at 2.5 mm, residual = the ellipsoid scene plus seeded noise at
    30 dB), all scans in one launch with the autotuned candidate chunk, and the top 256
    per scan. Also times the same correlations as one launch per scan."""
    import numpy as np
    import torch
    from threadpoolctl import threadpool_limits

    from dsplat import FMCWModel, Scan, workloads
    from dsplat.certificate import certificate_scan
    from dsplat.groundtruth import planar_array
    from dsplat.kernels import surfel

    scene = workloads.ellipsoid_scene()
    cfg = scene["config"]
    dev = torch.device("cuda", 0)
    sc, sn, sw = (scene[k].to(dev) for k in ("centers", "normals", "weights"))
    centers, normals = (x.to(dev) for x in workloads.certificate_candidates())
    C, K = len(centers), cfg.num_range_bins
    candidates = dict(centers=centers, normals=normals, areas=torch.full((C,), float(scene["areas"][0]), device=dev))
    models, residuals = [], []
    for i, side in enumerate((64, 96, 128)):
        pos = planar_array(side, workloads.PITCH).to(dev)
        data = surfel.render(cfg, pos, sc, sn, sw, 0.0)
        sigma = float(data.abs().square().mean().sqrt()) * 10 ** (-30 / 20)
        noise = workloads.complex_normal(tuple(data.shape), torch.Generator().manual_seed(40 + i)).to(dev)
        models.append(FMCWModel(Scan(data + sigma * noise, pos, cfg)))
        residuals.append(models[-1].scan.data)
    A = sum(m.apertures for m in models)
    evals = A * K * C
    stream = torch.cuda.current_stream(dev)
    scans = [(m.scan.scan_positions_m, r) for m, r in zip(models, residuals)]

    def step():
        return certificate_scan(models, residuals, candidates, "complex", top=256)

    for _ in range(max(1, args.warmup)):  # the first call autotunes the chunk
        step()
    torch.cuda.synchronize()
    from dsplat import autotune

    tuning = autotune.record("scan_batched")
    steps = max(1, min(args.steps, 10))
    t0 = time.perf_counter()
    for _ in range(steps):
        scores, ids = step()
    torch.cuda.synchronize()
    ms_per_step = (time.perf_counter() - t0) * 1e3 / steps
    kernel_ms, _ = timed(lambda: surfel.scan_batched(cfg, scans, centers, normals, 0.0), 3, stream)
    single_ms, _ = timed(lambda: [surfel.correlate(cfg, p, centers, normals, r, 0.0) for p, r in scans], 3, stream)
    sample_domain = cfg.splat_form == "samples" and cfg.adc_samples == 128
    scan_flop = TD_SCAN_FLOP if sample_domain else SCAN_FLOP
    achieved = scan_flop * evals / (kernel_ms * 1e-3) / 1e12
    # the top candidates lie near the surface: distance of the best per scan to the nearest surfel
    best = centers[ids[:, :16].flatten()]
    near = torch.cdist(best, sc).min(1).values.reshape(3, 16).median(1).values
    result = {
        "metric": "batched certificate scan evals/s (aperture·bin·candidate)",
        "value": evals / (ms_per_step * 1e-3),
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": max(1, args.warmup),
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp16 hi+lo x3 MFMA, fp32 accumulate, f64 pair phase, f64 range reduction" if sample_domain
        else "f32 (f64 pair phase, f64 block reduction)",
        "data": "synthetic",
        "config": {"workload": "certificate scan: 262144 candidates (64^3 grid, 70 mm box) x scans of 64^2, 96^2 "
                   "and 128^2 apertures @ 2.5 mm, 128 range bins, top 256 per scan", "apertures": A,
                   "bins": K, "candidates": C, "evals_per_step": evals, "parallelism": "single GPU"},
        "kernels_ms": {"scan_batched": kernel_ms, "one_launch_per_scan": single_ms},
        "autotune": tuning,
        "median_distance_top16_to_surface_m": [float(x) for x in near],
        "roofline": {"kernel": "scan_td_kernel<false>" if sample_domain else "scan_kernel<2>",
                     "bound": "valu-issue" if sample_domain else "valu", "achieved": achieved,
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": None, "flop_per_eval": scan_flop,
                     "peak_basis": "FP32 dense (FP32-accurate; FP16 hi+lo x3 MFMA)" if sample_domain
                     else "FP32 vector"},
    }
    if not args.no_cpu:
        from oracle import surfel as O

        pos = models[0].scan.scan_positions_m.cpu().numpy()
        res = residuals[0].cpu().numpy()
        cc, cn = centers.cpu().numpy(), normals.cpu().numpy()
        with threadpool_limits(limits=1):
            t = time.perf_counter()
            O.correlate(pos, cc[:8], cn[:8], res, cfg, 0.0)
            probe = time.perf_counter() - t
            count = int(min(1024, max(8, 8 * args.cpu_seconds / max(probe, 1e-3))))
            ids = np.linspace(0, C - 1, count).astype(np.int64)
            t = time.perf_counter()
            O.correlate(pos, cc[ids], cn[ids], res, cfg, 0.0)
            seconds = time.perf_counter() - t
        sample = len(pos) * K * count
        result["cpu_baseline"] = {
            "value": sample / seconds, "unit": "evals/s", "cores": 1, "kind": "port",
            "sample": f"float64 numpy oracle correlate of {count} of {C} candidates against the 64^2 scan "
            f"({sample:.3e} evals, {seconds:.1f} s, 1 thread)",
        }  # fmt: skip
    return result


DSFW_FORMS = {
    # plain FP32: bin-domain splat kernels, FP32 SYRK normal equations, candidates scored in
    # chunks, host-loop quasi-Newton search, backtracking and elimination
    "plain_fp32": ("bins", dict()),
    # mixed-precision batched: sample-domain MFMA splat, the projection writing split BF16
    # blocks for the split SYRK (hi hi + hi lo + lo hi), one-launch orientation choice,
    # quasi-Newton search and elimination, one design block per normal-equation pass (2^32 elements)
    "mixed_batched": ("samples", dict(gram="fp16x3", batched=True, dictionary_elements=2**32, row_chunk=2**24)),
}


def _relative_deviation(a, b):
    return max(abs(x - y) / max(abs(x), 1e-30) for x, y in zip(a, b, strict=True))


def dsfw_compare(steps, warmup=1, adc_samples=256, truth_draws=4000, front_count=None, controls=("mixed_batched",),
                 same_state_every=5, only=None):
    """The full DSFW fit in the plain FP32 and the mixed-precision batched forms from the same
    problem and initial state: `warmup` untimed steps, then `steps` steps each timed between
    device synchronizations (median and mean), the NMSE after every step, the final median
    centre error to the truth, and two parity figures.

    * Trajectory deviation: max over steps of |nmse_mixed - nmse_plain| / nmse_plain. The
      search takes greedy decisions on near-tied scores, so any rounding-level change of the
      inputs moves it; `controls` reruns the named forms with the target perturbed by complex
      Gaussian noise of 1e-7 of its RMS (the float32 rounding level), whose deviation from the
      unperturbed run of the same form measures that sensitivity.
    * Same-state agreement: every `same_state_every`-th geometry of the plain run is fitted
      again (VarPro: normal equations, solve, rendered objective) by both forms; the largest
      relative difference of the two NMSEs is the mixed form's evaluation error on identical
      states."""
    import statistics
    from dataclasses import replace

    import torch

    from dsplat import DSFWConfig, FMCWModel, Scan, workloads
    from dsplat.dsfw import initialize, step
    from dsplat.spatial import SpatialSearchConfig
    from dsplat.varpro import varpro

    problem = workloads.dsfw_problem(adc_samples=adc_samples, truth_draws=truth_draws, front_count=front_count)
    target, positions = problem["target"], problem["model"].scan.scan_positions_m
    chirp = problem["model"].scan.fmcw_config
    S = len(problem["initial"]["centers"])
    noise = workloads.complex_normal(tuple(target.shape), torch.Generator().manual_seed(77)).to(target.device)
    perturbed = target + noise * (float(target.abs().square().mean().sqrt()) * 1e-7)
    search = SpatialSearchConfig()

    def setup(name, data):
        form, extra = DSFW_FORMS[name]
        model = FMCWModel(Scan(data, positions, replace(chirp, splat_form=form)))
        return model, DSFWConfig(batch=256, iterations=steps, **extra)

    def run(name, data, keep_states=False):
        model, config = setup(name, data)
        state = initialize(model, data, problem["initial"], config, parameters=problem["parameters"])
        for _ in range(warmup):
            state = step(state, config, search)
        torch.cuda.synchronize()
        times, states = [], [state.params] if keep_states else []
        for i in range(steps):
            t0 = time.perf_counter()
            state = step(state, config, search)
            torch.cuda.synchronize()
            times.append((time.perf_counter() - t0) * 1e3)
            if keep_states:
                states.append(state.params)
            if (i + 1) % 25 == 0:
                print(f"[dsfw] {name}{' (perturbed)' if data is perturbed else ''} step {i + 1}/{steps}: "
                      f"nmse {state.history[-1]['nmse']:.5f}, median {statistics.median(times):.1f} ms", file=sys.stderr,
                      flush=True)  # fmt: skip
        error = torch.cdist(state.params["centers"], problem["truth"]["centers"]).min(1).values
        objective = [h["objective"] for h in state.history]
        record = {
            "ms_per_step_median": statistics.median(times),
            "ms_per_step_mean": statistics.fmean(times),
            "ms": times,
            "nmse": [h["nmse"] for h in state.history],
            "final_nmse": state.history[-1]["nmse"],
            "median_center_error_m": float(error.median()),
            "objective_non_increasing": all(b <= a for a, b in zip(objective, objective[1:])),
            "steps_with_geometry_update": sum(1 for h in state.history[1:] if h.get("lm_accepted")),
        }
        return record, states

    if only is not None:  # one form's steps alone (for kernel traces of that form)
        record, _ = run(only, target)
        del record["ms"]
        return {"steps": steps, "warmup": warmup, "surfels": S, "bins": adc_samples, "forms": {only: record}}
    plain, states = run("plain_fp32", target, keep_states=True)
    mixed, _ = run("mixed_batched", target)
    out = {"plain_fp32": plain, "mixed_batched": mixed}
    control = {}
    for name in controls:
        record, _ = run(name, perturbed)
        control[name] = _relative_deviation(out[name]["nmse"], record["nmse"])
        out[name + "_perturbed"] = record
    # identical geometries, fitted by both forms on the models the fits use (range subspace
    # of the parameter box, as initialize prepares them)
    fits = {}
    for name in DSFW_FORMS:
        model, config = setup(name, target)
        fits[name] = (model.prepare(problem["parameters"]), config)
    power = float(target.abs().square().mean())
    same = []
    for params in states[::same_state_every] + states[-1:]:
        nmse = [varpro(model, params, target, config).objective / power for model, config in fits.values()]
        same.append(abs(nmse[1] - nmse[0]) / nmse[0])
    for record in out.values():
        del record["ms"]
    return {
        "workload": f"DSFW fit: {S} truth surfels ({'front faces of %d ellipsoid draws' % truth_draws if front_count is None else '%d front faces' % front_count}), "
        f"128x128 array @ 2.5 mm, {adc_samples} ADC samples ({adc_samples} bins), float64 synthesis + 30 dB noise, "
        f"random initial surfels, batch 256, spatial search, joint Gauss-Newton",
        "surfels": S, "bins": adc_samples, "steps": steps, "warmup": warmup,
        "speedup_median": plain["ms_per_step_median"] / mixed["ms_per_step_median"],
        "speedup_mean": plain["ms_per_step_mean"] / mixed["ms_per_step_mean"],
        "nmse_max_relative_deviation": _relative_deviation(plain["nmse"], mixed["nmse"]),
        "control_nmse_max_relative_deviation": control,
        "control_note": "same form, target + 1e-7 RMS complex noise: the search's own sensitivity",
        "same_state_nmse_max_relative_deviation": max(same),
        "same_state_nmse_relative_deviation": same,
        "same_state_count": len(same),
        "forms": out,
    }  # fmt: skip


def dsfw_workload(args):
    """The DSFW comparison on its own (see dsfw_compare); both forms get a perturbed control."""
    draws = args.truth_draws or 4000
    only = None if args.dsfw_form == "both" else args.dsfw_form
    result = dsfw_compare(args.steps, max(1, args.warmup), args.adc_samples, draws,
                          front_count=args.front_count or None, controls=("plain_fp32", "mixed_batched"),
                          only=only)  # fmt: skip
    if only is not None:
        record = result["forms"][only]
        return {"metric": f"DSFW step ms (median), {only} form", "value": record["ms_per_step_median"], "unit": "ms",
                "n_gpus": 1, "steps": result["steps"], "warmup": result["warmup"], "higher_is_better": False,
                "data": "synthetic", "dsfw": result}  # fmt: skip
    mixed = result["forms"]["mixed_batched"]
    return {
        "metric": "DSFW step ms (median), mixed-precision batched form",
        "value": mixed["ms_per_step_median"],
        "unit": "ms",
        "n_gpus": 1,
        "steps": result["steps"],
        "warmup": result["warmup"],
        "ms_per_step": mixed["ms_per_step_median"],
        "higher_is_better": False,
        "dtype": "mixed form: fp16 hi+lo x3 MFMA split-operand normal equations and scan, fp32 MFMA splat at 256 "
        "samples, f32 Cholesky first with the f64 shifted fallback",
        "data": "synthetic",
        "config": {"workload": result["workload"], "surfels": result["surfels"], "bins": result["bins"],
                   "parallelism": "single GPU"},
        "dsfw": result,
    }


SYRK_SHAPES = ((118784, 2048), (1048576, 2035))  # (rows, columns): 29 rows per aperture of 128^2 and 4096


def syrk_workload(args):
    """The split-operand SYRK of the mixed DSFW normal equations (upper triangle of A^T A,
    FP16 hi + lo operands in 32-row column tiles, column scales 2^e_s, three products on the
    FP16 matrix cores, float64 reduction), on seeded Gaussian blocks of 118784 x 2048 and
    1048576 x 2035; beside it the FP32 SYRK of the same float32 rows. Executed work: 3 products
    of m n (n + 1) FLOP over the upper tiles."""
    import torch

    from dsplat.kernels import blas

    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream(dev)
    shapes = {}
    for m, n in SYRK_SHAPES:
        rows = torch.randn(m, n, device=dev, generator=torch.Generator(device=dev).manual_seed(1))
        exponents = torch.floor(torch.log2(2.0**14 / rows.abs().amax(0))).float()
        block = blas.SplitBlock.of(rows, "fp16", exponents)
        out = torch.zeros(n, n, device=dev)
        split_ms, _ = timed(lambda block=block, out=out: blas.syrk_split(block, out, None), max(1, args.steps), stream)
        fp32_ms, _ = timed(lambda rows=rows, out=out: blas.syrk(rows, out, "fp32"), 2, stream)
        flop = float(m) * n * (n + 1)  # one product of the upper triangle
        shapes[f"{m}x{n}"] = {
            "split_fp16_ms": split_ms,
            "split_fp16_executed_tflops": 3 * flop / split_ms / 1e9,
            "split_fp16_frac_of_fp16_peak": 3 * flop / split_ms / 1e9 / 2516.6,
            "fp32_ms": fp32_ms,
            "fp32_frac_of_fp32_peak": flop / fp32_ms / 1e9 / 157.3,
        }
        del rows, block, out
        torch.cuda.empty_cache()
    big = shapes[f"{SYRK_SHAPES[1][0]}x{SYRK_SHAPES[1][1]}"]
    return {
        "metric": "split-operand SYRK ms (1048576 x 2035, FP16 hi+lo x3)",
        "value": big["split_fp16_ms"],
        "unit": "ms",
        "n_gpus": 1,
        "higher_is_better": False,
        "dtype": "fp16 hi+lo x3 MFMA, fp32 accumulate, f64 split-K reduction",
        "data": "synthetic",
        "config": {"workload": "upper-triangle A^T A of seeded Gaussian float32 blocks, pre-split FP16 hi/lo",
                   "shapes": [list(x) for x in SYRK_SHAPES]},
        "shapes": shapes,
        "roofline": {"kernel": "syrk_split_kernel", "bound": "mfma",
                     "achieved": big["split_fp16_executed_tflops"], "peak": 2516.6, "unit": "TFLOP/s",
                     "frac": big["split_fp16_frac_of_fp16_peak"], "traffic": None},
    }


def valu_workload(args):
    """The VALU issue rate: eight independent FMA chains per lane on every SIMD, FP32 and FP64
    (dsplat_calibrate_valu), as wave64 instructions per second and FLOP/s (2 per lane FMA)."""
    import torch

    from dsplat import _lib

    dev = torch.device("cuda", 0)
    blocks, iters = 256 * 16, 4096  # 16 workgroups of 4 waves per CU
    stream = torch.cuda.current_stream(dev)
    rates = {}
    for name, fp64 in (("fp32", 0), ("fp64", 1)):
        out = torch.empty(blocks * 256, dtype=torch.float64 if fp64 else torch.float32, device=dev)

        def launch():
            _lib.call("dsplat_calibrate_valu", fp64, blocks, iters, _lib.ptr(out), _lib.stream(dev))

        for _ in range(max(1, args.warmup)):
            launch()
        ms, _ = timed(launch, max(3, args.steps), stream)
        instructions = blocks * 4 * 8 * iters  # wave64 FMA instructions
        rates[name] = {"ms": ms, "wave_instructions_per_s": instructions / (ms * 1e-3),
                       "tflops": instructions * 64 * 2 / (ms * 1e-3) / 1e12,
                       "cycles_per_instruction_per_simd_at_2_4_ghz": SIMDS * CLOCK_GHZ * 1e9 / (instructions / (ms * 1e-3))}
    return {
        "metric": "VALU FMA issue rate (wave64 instructions/s), FP32",
        "value": rates["fp32"]["wave_instructions_per_s"],
        "unit": "wave64 instructions/s",
        "n_gpus": 1,
        "steps": max(3, args.steps),
        "higher_is_better": True,
        "dtype": "f32 and f64",
        "data": "synthetic",
        "config": {"workload": f"{blocks} workgroups x 256 lanes x 8 independent FMA chains x {iters} steps"},
        "rates": rates,
        "published_fp32_peak_tflops": PEAK_FP32_TFLOPS,
    }


def copy_workload(args):
    """A plain device-to-device copy of 2^28 float32 (1 GiB read, 1 GiB written) per step:
    the known-size kernel that scales the FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch

    dev = torch.device("cuda", 0)
    src = torch.ones(1 << 28, device=dev)
    dst = torch.empty_like(src)
    stream = torch.cuda.current_stream(dev)
    for _ in range(max(1, args.warmup)):
        dst.copy_(src)
    ms, _ = timed(lambda: dst.copy_(src), max(1, args.steps), stream)
    nbytes = 2 * src.numel() * src.element_size()
    return {
        "metric": "device copy GB/s (read + write)",
        "value": nbytes / (ms * 1e-3) / 1e9,
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": max(1, args.steps),
        "ms_per_step": ms,
        "higher_is_better": True,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "copy of 2^28 float32 (1 GiB each way)", "bytes_read": nbytes // 2,
                   "bytes_written": nbytes // 2},
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(
            f"bench.py runs on one GPU: --gpus {args.gpus} is not supported in this build "
            "(the multi-GPU sharding is described, not built).",
            file=sys.stderr,
        )
        sys.exit(2)
    if args.workload == "gram":
        print(json.dumps(gram_workload(args)), flush=True)
        return
    if args.workload == "scan":
        print(json.dumps(scan_workload(args)), flush=True)
        return
    if args.workload == "dsfw":
        print(json.dumps(dsfw_workload(args)), flush=True)
        return
    if args.workload == "synth":
        print(json.dumps(synth_workload(args)), flush=True)
        return
    if args.workload == "syrk":
        print(json.dumps(syrk_workload(args)), flush=True)
        return
    if args.workload == "copy":
        print(json.dumps(copy_workload(args)), flush=True)
        return
    if args.workload == "valu":
        print(json.dumps(valu_workload(args)), flush=True)
        return
    import torch

    from dsplat import workloads
    from dsplat.kernels import surfel

    work = workloads.ellipsoid_scene()
    cfg = work["config"]
    dev = torch.device("cuda", 0)
    positions, centers, normals = (work[k].to(dev) for k in ("positions", "centers", "normals"))
    weights, up = work["weights"].to(dev), work["cotangent"].to(dev)
    A, K, S = len(positions), cfg.num_range_bins, len(centers)
    evals = A * K * S

    def step():
        out = surfel.render(cfg, positions, centers, normals, weights, 0.0)
        grads = surfel.vjp(cfg, positions, centers, normals, weights, up, 0.0)
        return out, grads

    # Clock ramp: the GPU reaches its sustained clocks only after tens of milliseconds of
    # work, so untimed steps run for args.ramp_ms of wall time before the W warm-up steps.
    ramp_steps = 0
    torch.cuda.synchronize()
    t_ramp = time.perf_counter()
    while (time.perf_counter() - t_ramp) * 1e3 < args.ramp_ms:
        step()
        ramp_steps += 1
        if ramp_steps % 16 == 0:
            torch.cuda.synchronize()
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    stream = torch.cuda.current_stream(dev)
    marks = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for m in marks:
        m[0].record(stream)
        surfel.render(cfg, positions, centers, normals, weights, 0.0)
        m[1].record(stream)
        surfel.vjp(cfg, positions, centers, normals, weights, up, 0.0)
        m[2].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    fwd_ms = sum(m[0].elapsed_time(m[1]) for m in marks) / args.steps
    vjp_ms = sum(m[1].elapsed_time(m[2]) for m in marks) / args.steps
    ms_per_step = wall * 1e3 / args.steps

    # the forward-mode product (not part of the metric), timed after the steps
    dc, dn, dw = (work[k].to(dev) for k in ("dcenters", "dnormals", "dweights"))
    jvp_ms, _ = timed(lambda: surfel.jvp(cfg, positions, centers, normals, weights, dc, dn, dw, 0.0), 5, stream)

    sample_domain = cfg.splat_form == "samples" and cfg.adc_samples in (128, 256)
    moments = sample_domain and cfg.pad_factor >= 2
    groups = {  # the launches each timed call makes
        "render": ["td_render_h16_kernel<false, false>"] if sample_domain else ["render_kernel<2, 8>"],
        "vjp": ["vjp_moments_prep", "vjp_moments_kernel<false>", "vjp_reduce"] if moments
        else (["vjp_prep_kernel", "td_vjp_kernel<8, false>", "vjp_reduce"] if sample_domain
              else ["vjp_kernel<2>", "vjp_reduce"]),
    }  # fmt: skip
    pairs = A * S
    valu_rate = valu_workload(argparse.Namespace(warmup=1, steps=3))["rates"] if sample_domain else None

    def valu_issue(kernel, ms, over):
        """The kernel's VALU instruction count against the issue rates, over `ms` of the
        launch set named by `over` (the kernel alone, or its whole launch group)."""
        if kernel not in VALU_PER_LAUNCH:
            return None
        count = VALU_PER_LAUNCH[kernel]
        out = {"kernel": kernel, "instructions_per_launch": count, "instructions_per_64_pairs": count * 64 / pairs,
               "time_basis": over, "ms": ms,
               "source": f"{VALU_PROFILE.relative_to(ROOT)} (rocprofv3 --pmc SQ_INSTS_VALU)",
               "issue_ms_at_published_peak": count / VALU_PEAK_PER_S * 1e3,
               "frac_of_published_peak": count / VALU_PEAK_PER_S * 1e3 / ms,
               "basis": "published FP32 peak: one wave64 FP32 instruction per 2 cycles on each of 1024 SIMDs "
                        "at 2.4 GHz"}  # fmt: skip
        if valu_rate is not None:
            measured = valu_rate["fp32"]["wave_instructions_per_s"]
            out["measured_fma_rate"] = measured
            out["frac_of_measured_rate"] = count / measured * 1e3 / ms
        return out

    if sample_domain:
        # forward: the real multiply-adds of the symmetric sample sums per (a, n, s)
        name, ms = "render", fwd_ms
        achieved = TD_FWD_FLOP * evals / (ms * 1e-3) / 1e12
        roofline = {
            "kernel": name,
            "dominant": fwd_ms >= vjp_ms,  # the render is the longer launch group of the step
            "launches": groups[name],
            "bound": "valu-issue",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": measured_traffic(groups[name]),
            "traffic_source": f"{TRAFFIC.relative_to(ROOT)} (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE per dispatch, "
            "scaled by the same counters on a 1 GiB copy)",
            "flop_per_eval": TD_FWD_FLOP,
            "peak_basis": "FP32 peak: results are FP32-accurate, executed as FP16 hi+lo x3 MFMA products",
            "cotangent_form_count": {"flop_per_eval": FWD_FLOP, "tflops": FWD_FLOP * evals / (ms * 1e-3) / 1e12,
                                     "frac": FWD_FLOP * evals / (ms * 1e-3) / 1e12 / PEAK_FP32_TFLOPS},
            "valu_issue": valu_issue(groups[name][0], ms, "the render launch"),
            "executed_fp16_mfma": {
                "tflops": TD_EXEC_FLOP_PER_PAIR * pairs / (ms * 1e-3) / 1e12,
                "frac_of_fp16_peak": TD_EXEC_FLOP_PER_PAIR * pairs / (ms * 1e-3) / 1e12 / PEAK_FP16_TFLOPS,
            },
            # the VJP by bin moments: no per-(a, n, s) sums; its bound is the per-pair VALU work
            "vjp": {"launches": groups["vjp"], "ms": vjp_ms, "traffic": measured_traffic(groups["vjp"]),
                    "valu_issue": valu_issue(groups["vjp"][1], vjp_ms, "the VJP launch group: pre-pass, moments "
                                             "kernel and reduction") if moments else None},
        }  # fmt: skip
    else:
        kernels = {"render": (fwd_ms, FWD_FLOP * evals), "vjp": (vjp_ms, VJP_FLOP * evals)}
        name = max(kernels, key=lambda k: kernels[k][0])
        ms, flop = kernels[name]
        achieved = flop / (ms * 1e-3) / 1e12
        roofline = {"kernel": name, "launches": groups[name], "bound": "valu", "achieved": achieved,
                    "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                    "traffic": measured_traffic(groups[name]), "flop_per_eval": FWD_FLOP if name == "render" else VJP_FLOP,
                    "peak_basis": "FP32 vector"}  # fmt: skip
    result = {
        "metric": "fwd+VJP splat evals/s (aperture·bin·surfel)",
        "value": evals / (ms_per_step * 1e-3),
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ramp_steps": ramp_steps,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp16 hi+lo x3 MFMA, fp32 accumulate, f64 pair phase" if sample_domain
        else "f32 (f64 pair phase)",
        "data": "synthetic",
        "config": {
            "workload": "FMCW surfel splat fwd+VJP: 128x128 apertures @ 2.5 mm, 128 range bins, "
            "2048 ellipsoid surfels",
            "apertures": A,
            "bins": K,
            "surfels": S,
            "evals_per_step": evals,
            "parallelism": "single GPU",
        },
        "kernels_ms": {"render": fwd_ms, "vjp": vjp_ms, "jvp": jvp_ms},
        "splat_form": "sample-domain MFMA" if sample_domain else "bin-domain VALU",
        "roofline": roofline,
    }
    del work["dcenters"], work["dnormals"], work["dweights"]
    torch.cuda.empty_cache()
    if args.dsfw_steps > 0:
        print(f"[bench] splat: {result['value']:.4g} evals/s; DSFW comparison, {args.dsfw_steps} steps per form",
              file=sys.stderr, flush=True)  # fmt: skip
        result["dsfw"] = dsfw_compare(args.dsfw_steps, 1, args.adc_samples, args.truth_draws,
                                      front_count=args.front_count or None)  # fmt: skip
    if not args.no_cpu and int(os.environ.get("RANK", "0")) == 0:
        print("[bench] CPU baseline", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_torch_fp32(work)
        result["cpu_float64_oracle"] = cpu_baseline(work, args.cpu_seconds)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
