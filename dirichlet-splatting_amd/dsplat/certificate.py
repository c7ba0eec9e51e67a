"""Residual certificate: normalized correlations, candidate grids, peaks and refinement.

Same functions as dsplat/certificate.py of the reference (certificate.py:11-201):
the certificate |s|^2 / (|a|^2 |r|^2) ranks insertions (s = Re(a^H r) for real
coefficients), candidate grids and greedy separated peaks select them, and a batched
per-candidate quasi-Newton ascent with backtracking refines them.
"""

from dataclasses import dataclass
from math import ceil

import torch

from .kernels import search

PEAK_POOL = 64  # candidates kept per wanted peak before suppression


def certificate(model, params, residual, mode, *, gradient=False, normals=False):
    """Squared normalized certificate per atom, and with gradient its derivative by
    centers (and normals); per-candidate algebra in float64 keeps the 1e-60 norm floor
    representable (certificate.py:11-31)."""
    if gradient:
        c, n, dc, dn = model.correlate_gradient(params, residual, include_normals=normals)
    else:
        c, n = model.correlate(params, residual)
    c, n = c.to(torch.complex128), n.double().clamp(min=1e-60)
    power = residual.abs().square().sum().double().clamp(min=1e-30)
    s = c.real if mode == "real" else c
    value = s.abs().square() / n / power
    if not gradient:
        return value.float()
    dc = dc.to(torch.complex128)
    dn = torch.where((n > 1e-60)[:, None], dn.double(), 0)
    ds = 2 * s[:, None] * dc.real if mode == "real" else 2 * (c.conj()[:, None] * dc).real
    return value.float(), ((ds - value[:, None] * power * dn) / n[:, None] / power).float()


@dataclass(frozen=True)
class CertificateSearchConfig:
    peaks: int = 1024  # separated coarse peaks refined per search
    separation: float = 0.0005  # center units, for coarse peak selection
    iterations: int = 40  # quasi-Newton ascent steps
    trust_radius: float = 1.0  # times each parameter's scale
    backtracks: int = 8


QN_MIN_BATCH = 256  # candidates per one-launch quasi-Newton search (one workgroup each)


def candidate_grid(lower, upper, spacing, device="cuda"):
    """Regular grid including both box corners; the actual pitch never exceeds spacing."""
    axes = [
        torch.linspace(a, b, 1 + ceil((b - a) / spacing), device=device) for a, b in zip(lower, upper, strict=True)
    ]
    return torch.stack(torch.meshgrid(*axes, indexing="ij"), -1).reshape(-1, len(axes))


def select_peaks(scores, centers, count, separation):
    """Greedy nonmaximum suppression: best-first indices at least separation apart.

    Without separation it is a stable sort. Otherwise, when the pool of the top
    PEAK_POOL * count scores already yields every wanted peak, the rest cannot change
    the answer (everything at or above a level is decided before anything below it),
    so the one-workgroup kernel runs on the pool (certificate.py:52-70).
    """
    if separation == 0:
        order = scores.argsort(descending=True, stable=True)[:count]
        return order[torch.isfinite(scores[order])]
    wanted = min(count, len(scores))
    if len(scores) > PEAK_POOL * wanted:
        level = scores.topk(PEAK_POOL * wanted).values[-1]
        pool = (scores >= level).nonzero().flatten()
        chosen = search.peaks(scores[pool], centers[pool], wanted, separation)
        if len(chosen) == wanted:
            return pool[chosen]
    return search.peaks(scores, centers, count, separation)


def fibonacci_sphere(count, device="cuda"):
    """Nearly uniform unit vectors [count, 3]."""
    j = torch.arange(count, device=device, dtype=torch.float32)
    z = 1 - 2 * (j + 0.5) / count
    angle = j * (torch.pi * (3 - 5**0.5))
    radius = (1 - z.square()).sqrt()
    return torch.stack((radius * angle.cos(), radius * angle.sin(), z), -1)


@torch.no_grad()
def orient_candidates(model, params, residual, count, mode, table=None, chunk=8192, *, batched=False):
    """Give every position the best of its own normal and a count-normal sphere bank;
    returns the candidates with the chosen normals and their certificates. batched scores
    every candidate and picks its normal in one launch (no [C, 1 + count] sums)."""
    centers = params["centers"]
    bank = fibonacci_sphere(count, centers.device)
    power = residual.abs().square().sum().double().clamp(min=1e-30)
    if batched and hasattr(model, "orient_select"):
        value, ids = model.orient_select(params, residual, bank, table, mode == "real")
        picked = bank[(ids - 1).clamp(min=0)]
        normals = torch.where((ids > 0)[:, None], picked, params["normals"])
        return {**params, "normals": normals}, (value / power).float()
    scores, normals = [], []
    for start in range(0, len(centers), chunk):
        block = {key: value[start : start + chunk] for key, value in params.items()}
        c, n = model.correlate_orientations(block, residual, bank, table)
        c, n = c.to(torch.complex128), n.double()
        s = c.real if mode == "real" else c
        best, ids = (s.abs().square() / n.clamp(min=1e-60) / power).max(1)
        chosen = block["normals"].clone()
        banked = ids > 0
        chosen[banked] = bank[ids[banked] - 1]
        scores.append(best.float())
        normals.append(chosen)
    return {**params, "normals": torch.cat(normals)}, torch.cat(scores)


def _controls(reference, center_spec, normal_spec, masks):
    """[4, 6] per-coordinate scale, trainable mask, lower and upper bounds."""
    controls = reference.new_zeros((4, 6))
    controls[0, :3] = center_spec.scale
    controls[1, :3] = masks["centers"]
    controls[2, :3] = reference.new_tensor(-torch.inf if center_spec.lower is None else center_spec.lower)
    controls[3, :3] = reference.new_tensor(torch.inf if center_spec.upper is None else center_spec.upper)
    if normal_spec is not None:
        controls[0, 3:], controls[1, 3:] = normal_spec.scale, 1
    return controls


@torch.no_grad()
def search_certificate(
    model, residual, candidates, center_spec, config, mode, *, normal_spec=None,
    initial_scores=None, full_model=None, full_residual=None, batched=False,
):  # fmt: skip
    """Refine separated peaks in xyz (and unit normals) by batched quasi-Newton ascent
    with damped BFGS and per-candidate backtracking; coarse peaks stay beside the refined
    ones and all are reranked on full_model when given (certificate.py:106-201). batched
    runs the whole ascent of every candidate in one launch instead of a host loop."""
    reference = candidates["centers"]
    feasible = dict(candidates, centers=center_spec.project(reference, reference))
    specs = {"centers": center_spec}
    if normal_spec is not None:
        specs["normals"] = normal_spec
        feasible["normals"] = normal_spec.project(candidates["normals"], candidates["normals"])
    masks = {key: spec.mask(feasible[key]) for key, spec in specs.items()}
    controls = _controls(reference, center_spec, normal_spec, masks)

    def ascent(params, gradient):
        parts = [
            spec.tangent(gradient[:, 3 * i : 3 * i + 3], params[key]) * masks[key] * spec.scale
            for i, (key, spec) in enumerate(specs.items())
        ]
        return -torch.cat(parts, -1)

    def score(params, gradient=False):
        return certificate(model, params, residual, mode, gradient=gradient, normals="normals" in specs)

    coarse_scores = score(feasible) if initial_scores is None else initial_scores
    ids = select_peaks(coarse_scores, feasible["centers"], config.peaks, config.separation)
    coarse = {key: value[ids] for key, value in feasible.items()}
    # One workgroup per candidate fills the GPU only with enough candidates; fewer (the
    # full-aperture refinement of a few dozen peaks) run the host loop, whose every launch
    # spreads all apertures over the chip.
    if batched and hasattr(model, "qn_search") and not center_spec.normalize and len(ids) >= QN_MIN_BATCH:
        spec = _qn_spec(controls, center_spec, normal_spec, masks, config, mode, residual)
        centers, normals, value = model.qn_search(coarse, residual, spec)
        return _rerank(coarse, dict(coarse, centers=centers, normals=normals), coarse_scores[ids], value, full_model,
                       full_residual, mode)  # fmt: skip
    params = {key: value.clone() for key, value in coarse.items()}
    value, grad = score(params, True)
    grad = ascent(params, grad)
    count, dim = grad.shape
    eye = torch.eye(dim, device=reference.device).expand(count, dim, dim)
    inverse = eye.clone()
    for _ in range(config.iterations):
        direction = search.direction(inverse, grad, params["normals"], controls[1], config.trust_radius)
        old = {key: v.clone() for key, v in params.items()}
        accepted = torch.zeros(count, dtype=torch.bool, device=reference.device)
        for backtrack in range(config.backtracks):
            centers, normals = search.trial(old["centers"], old["normals"], direction, controls, 0.5**backtrack)
            trial = dict(old, centers=centers, normals=normals)
            pending = torch.where(~accepted)[0]
            trial_value = score({key: v[pending] for key, v in trial.items()})
            keep = torch.zeros_like(accepted)
            keep[pending] = torch.isfinite(trial_value) & (trial_value > value[pending])
            for key in specs:
                params[key][keep] = trial[key][keep]
            accepted |= keep
            if accepted.all():
                break
        if not accepted.any():
            break
        new_value, new_grad = score(params, True)
        new_grad = ascent(params, new_grad)
        step = torch.cat([(params[key] - old[key]) / spec.scale for key, spec in specs.items()], -1)
        change = new_grad - grad
        sy = (step * change).sum(-1)
        valid = accepted & (sy > 1e-8 * step.norm(dim=-1) * change.norm(dim=-1)) & (sy > 1e-20)
        rho = torch.where(valid, 1 / sy.clamp(min=1e-30), 0)
        left = eye - rho[:, None, None] * step[:, :, None] * change[:, None, :]
        updated = left @ inverse @ left.transpose(-1, -2) + rho[:, None, None] * step[:, :, None] * step[:, None, :]
        inverse = torch.where(valid[:, None, None], updated, inverse)
        value, grad = new_value, new_grad
    return _rerank(coarse, params, coarse_scores[ids], value, full_model, full_residual, mode)


def _rerank(coarse, params, coarse_scores, value, full_model, full_residual, mode):
    """Coarse and refined candidates together, best first (on full_model when given)."""
    combined = {key: torch.cat((coarse[key], params[key])) for key in params}
    if full_model is None:
        scores = torch.cat((coarse_scores, value))
    else:
        scores = certificate(full_model, combined, full_residual, mode)
    order = scores.argsort(descending=True, stable=True)
    return {key: v[order] for key, v in combined.items()}, scores[order]


def _qn_spec(controls, center_spec, normal_spec, masks, config, mode, residual):
    """The constants of the one-launch ascent (dsplat_qn_spec)."""
    from ._lib import QnSpec

    ctl = controls.cpu()
    spec = QnSpec()
    spec.center_scale = float(center_spec.scale)
    spec.normal_scale = float(normal_spec.scale) if normal_spec is not None else 1.0
    spec.center_mask[:] = masks["centers"].float().cpu().tolist()
    spec.normal_mask[:] = masks["normals"].float().cpu().tolist() if normal_spec is not None else [1.0, 1.0, 1.0]
    spec.direction_mask[:] = ctl[1].tolist()
    spec.step_scale[:] = ctl[0].tolist()
    spec.lower[:] = ctl[2, :3].tolist()
    spec.upper[:] = ctl[3, :3].tolist()
    spec.width = 6 if normal_spec is not None else 3
    spec.normalize = int(bool(normal_spec.normalize)) if normal_spec is not None else 0
    spec.real_part = int(mode == "real")
    spec.iterations = int(config.iterations)
    spec.backtracks = int(config.backtracks)
    spec.radius = float(config.trust_radius)
    spec.power = float(residual.abs().square().sum().double().clamp(min=1e-30))
    return spec


@torch.no_grad()
def certificate_scan(models, residuals, candidates, mode="complex", top=256, chunk=None):
    """Certificates of every candidate against several scans of one chirp in one batched
    launch, and the `top` best per scan (stable: ties keep the lower index).

    models are FMCWModels sharing the chirp and material; residuals [A_i, K]. Returns
    (scores [n, C] float32, top ids [n, top] int64).
    """
    from .kernels.surfel import scan_batched

    config = models[0].scan.fmcw_config
    scans = [(m.scan.scan_positions_m, r) for m, r in zip(models, residuals, strict=True)]
    corr, norm = scan_batched(config, scans, candidates["centers"], candidates["normals"], models[0].fresnel_n, chunk)
    area = candidates["areas"].double()
    c = corr.to(torch.complex128) * area.sqrt()
    n = norm.double() * area
    power = torch.stack([r.abs().square().sum().double() for r in residuals]).clamp(min=1e-30)[:, None]
    s = c.real if mode == "real" else c
    scores = s.abs().square() / n.clamp(min=1e-60) / power
    ids = torch.stack([torch.argsort(row, descending=True, stable=True)[:top] for row in scores])
    return scores.float(), ids
