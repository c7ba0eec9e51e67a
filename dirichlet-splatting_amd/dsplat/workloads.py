"""Seeded synthetic workloads of the splat, Gram, certificate-scan and DSFW benchmarks.

Every draw uses a CPU torch.Generator with a fixed seed, so both machines build the
same inputs; tensors are returned on the CPU. Sizes:

  fmcw_small      32x32 array @ 2.5 mm, 200 surfels in a 60 mm cube at (0,0,0.35)
  ellipsoid_scene 128x128 array @ 2.5 mm, 2048 front-facing surfels on the
                  30/30/25 mm ellipsoid at (0,0,0.35), area = ellipsoid area / 4096
  certificate     262144 candidates on a 64^3 grid over the 70 mm box, scans of
                  64^2, 96^2 and 128^2 apertures
  dsfw_problem    4000 ellipsoid draws (front faces kept) or 2048 front faces, 128x128
                  array, float64 synthesis + 30 dB noise, random initial surfels
"""

import math

import torch

from .fmcw import FMCWConfig
from .groundtruth import planar_array

CENTER = (0.0, 0.0, 0.35)
SEMI_AXES = (0.030, 0.030, 0.025)
PITCH = 0.0025


def _gen(seed):
    return torch.Generator().manual_seed(seed)


def hemisphere_normals(count, generator):
    """normalize(N(0, I)) with n_z = -|n_z|: uniform on the hemisphere facing -z."""
    n = torch.randn(count, 3, generator=generator)
    n = torch.nn.functional.normalize(n, dim=1)
    n[:, 2] = -n[:, 2].abs()
    return n


def complex_normal(shape, generator):
    """Complex N(0, 1): each component has variance 1/2 (torch.randn convention)."""
    return torch.randn(*shape, dtype=torch.complex64, generator=generator)


def ellipsoid_area(a, b, c):
    """Surface area of the spheroid a = b > c (oblate closed form)."""
    if abs(a - b) > 1e-15 or c >= a:
        raise ValueError("closed form implemented for oblate spheroids a = b > c")
    e = math.sqrt(1 - c * c / (a * a))
    return 2 * math.pi * a * a * (1 + (1 - e * e) / e * math.atanh(e))


def ellipsoid_samples(count, seed, front_only=True, keep_all_draws=False):
    """Area-uniform points on the ellipsoid and outward unit normals.

    Directions uniform on the sphere are mapped to the ellipsoid and accepted with
    probability proportional to the area element (rejection sampling). With
    front_only, exactly `count` points with outward n_z < 0 are returned; with
    keep_all_draws, `count` accepted points are drawn and only the front ones kept.
    """
    a, b, c = SEMI_AXES
    g = _gen(seed)
    centers, normals = [], []
    have, drawn = 0, 0
    gmax = max(b * c, a * c, a * b)
    while (have < count) if not keep_all_draws else (drawn < count):
        v = torch.nn.functional.normalize(torch.randn(4 * count, 3, generator=g, dtype=torch.float64), dim=1)
        area = torch.sqrt((b * c * v[:, 0]) ** 2 + (a * c * v[:, 1]) ** 2 + (a * b * v[:, 2]) ** 2)
        accept = torch.rand(4 * count, generator=g, dtype=torch.float64) * gmax < area
        v = v[accept]
        if keep_all_draws:
            v = v[: count - drawn]
            drawn += len(v)
        x = v * torch.tensor([a, b, c], dtype=torch.float64)
        n = torch.nn.functional.normalize(x / torch.tensor([a * a, b * b, c * c], dtype=torch.float64), dim=1)
        if front_only:
            keep = n[:, 2] < 0
            x, n = x[keep], n[keep]
        centers.append(x + torch.tensor(CENTER, dtype=torch.float64))
        normals.append(n)
        have += len(x)
    centers, normals = torch.cat(centers), torch.cat(normals)
    if not keep_all_draws:
        centers, normals = centers[:count], normals[:count]
    return centers.float(), normals.float()


def fmcw_small(seed=0):
    """The CPU-runnable splat case: A = 1024, K = 128, S = 200."""
    g = _gen(seed)
    centers = torch.tensor(CENTER) + 0.03 * (2 * torch.rand(200, 3, generator=g) - 1)
    normals = hemisphere_normals(200, g)
    areas = torch.full((200,), 1e-5)
    weights = areas.sqrt().to(torch.complex64)
    return dict(
        config=FMCWConfig(),
        positions=planar_array(32, PITCH),
        centers=centers,
        normals=normals,
        areas=areas,
        weights=weights,
        cotangent=complex_normal((1024, 128), _gen(seed + 1)),
    )


def ellipsoid_scene(side=128, count=2048, seed=0):
    """The full-size render: 128^2 apertures, 2048 front-facing ellipsoid surfels."""
    centers, normals = ellipsoid_samples(count, seed)
    area = ellipsoid_area(*SEMI_AXES) / (2 * count)
    areas = torch.full((count,), area)
    g = _gen(seed + 2)
    return dict(
        config=FMCWConfig(),
        positions=planar_array(side, PITCH),
        centers=centers,
        normals=normals,
        areas=areas,
        weights=areas.sqrt().to(torch.complex64),
        cotangent=complex_normal((side * side, 128), _gen(seed + 1)),
        dcenters=1e-4 * torch.randn(count, 3, generator=g),
        dnormals=1e-2 * torch.randn(count, 3, generator=g),
        dweights=complex_normal((count,), g) * areas.sqrt(),
    )


def certificate_candidates(per_axis=64, seed=4):
    """64^3 grid over the 70 mm box around the scene, with random -z-facing normals."""
    axis = torch.linspace(-0.035, 0.035, per_axis)
    grid = torch.stack(torch.meshgrid(axis, axis, axis, indexing="ij"), -1).reshape(-1, 3)
    centers = grid + torch.tensor(CENTER)
    return centers, hemisphere_normals(len(centers), _gen(seed))


def dsfw_problem(adc_samples=256, truth_draws=4000, front_count=None, snr_db=30.0, device="cuda", side=128):
    """A DSFW fit: a float64 scan of ellipsoid surfels on the side x side array plus complex
    Gaussian noise (seed 5, sigma^2 = mean|y|^2 / 10^(snr/10)), and an initial guess of as
    many surfels uniform in the truth's box with 5 mm margin and random normals (seed 1).

    truth_draws area-uniform draws keep the front faces (about half; area = ellipsoid
    area / draws), or, with front_count, exactly that many front faces (area = ellipsoid
    area / (2 front_count)). Returns dict(model, target, initial, parameters, truth).
    """
    from .fmcw import FMCWModel, Scan
    from .groundtruth import synthesize
    from .parameters import ParameterConfig

    config = FMCWConfig(adc_samples=adc_samples)
    if front_count is None:
        centers, normals = ellipsoid_samples(truth_draws, 0, keep_all_draws=True)
        area = ellipsoid_area(*SEMI_AXES) / truth_draws
    else:
        centers, normals = ellipsoid_samples(front_count, 0)
        area = ellipsoid_area(*SEMI_AXES) / (2 * front_count)
    centers, normals = centers.to(device), normals.to(device)
    positions = planar_array(side, PITCH).to(device)
    amplitudes = torch.full((len(centers),), area**0.5, device=device)
    clean = synthesize(positions, centers, normals, amplitudes, config).data
    sigma = float(clean.abs().square().mean().sqrt()) * 10 ** (-snr_db / 20)
    target = clean + sigma * complex_normal(tuple(clean.shape), _gen(5)).to(device)
    g = _gen(1)
    low = centers.min(0).values.cpu() - 0.005
    high = centers.max(0).values.cpu() + 0.005
    count = len(centers)
    initial = dict(
        centers=(low + (high - low) * torch.rand(count, 3, generator=g)).to(device),
        normals=torch.nn.functional.normalize(torch.randn(count, 3, generator=g), dim=1).to(device),
        areas=torch.full((count,), area, device=device),
    )
    parameters = dict(
        centers=ParameterConfig(scale=5e-4, lower=tuple(low.tolist()), upper=tuple(high.tolist())),
        normals=ParameterConfig(scale=0.1, normalize=True),
        areas=ParameterConfig(trainable=False),
    )
    return dict(
        model=FMCWModel(Scan(target, positions, config)),
        target=target,
        initial=initial,
        parameters=parameters,
        truth=dict(centers=centers, normals=normals, area=area),
        noise_sigma=sigma,
    )
