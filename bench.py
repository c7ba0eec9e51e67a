"""Benchmark of the Dirichlet-splat hot path on one MI355X.

A step is one forward render plus one VJP of the full-size scene: 128x128 apertures
at 2.5 mm, 128 range bins (120 GHz, 15 GHz, 128 ADC samples, pad 2), 2048
front-facing surfels on the 30/30/25 mm ellipsoid at 0.35 m. The metric counts
(aperture, bin, surfel) evaluations per second of fwd+VJP; inputs are resident in
HBM before the timed region.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Prints one JSON line: value, ms_per_step, a roofline object for the dominant kernel
(HIP events on the launch stream) and the float64 oracle timed on this host's cores
as cpu_baseline.
"""

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "dirichlet-splatting_amd"))

# Algorithmic FP32 work per (a, k, s) evaluation: the cot-form kernel.
#   forward 13 FLOP: cot by angle addition (3 + 3), cot (1), complex accumulate (4 + 2)
#   VJP     19 FLOP: the forward's 13 plus the (1 + cot^2) u class sum (6)
FWD_FLOP, VJP_FLOP = 13, 19
PEAK_FP32_TFLOPS = 157.3  # MI355X published FP32 vector peak
PEAK_HBM_GBS = 8000.0


def parse():
    parser = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=20)
    parser.add_argument("--warmup", type=int, default=3)
    parser.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    parser.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU-baseline sample time")
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


def main():
    args = parse()
    if args.gpus != 1:
        print(
            f"bench.py runs on one GPU: --gpus {args.gpus} is not supported in this build "
            "(the multi-GPU sharding is described, not built).",
            file=sys.stderr,
        )
        sys.exit(2)
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

    kernels = {
        "render": (fwd_ms, FWD_FLOP * evals),
        "vjp": (vjp_ms, VJP_FLOP * evals),
    }
    name = max(kernels, key=lambda k: kernels[k][0])
    ms, flop = kernels[name]
    achieved = flop / (ms * 1e-3) / 1e12
    result = {
        "metric": "fwd+VJP splat evals/s (aperture·bin·surfel)",
        "value": evals / (ms_per_step * 1e-3),
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 per-pair phase)",
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
        "kernels_ms": {"render": fwd_ms, "vjp": vjp_ms},
        "roofline": {
            "kernel": name,
            "bound": "valu",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": None,
            "flop_per_eval": FWD_FLOP if name == "render" else VJP_FLOP,
        },
    }
    if not args.no_cpu and int(os.environ.get("RANK", "0")) == 0:
        result["cpu_baseline"] = cpu_baseline(work, args.cpu_seconds)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
