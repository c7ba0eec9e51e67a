"""The residual certificate and the candidate search built on it.

Public surface of the reference's dsplat/certificate.py:11-201 (certificate,
CertificateSearchConfig, candidate_grid, select_peaks, fibonacci_sphere, orient_candidates,
search_certificate). The certificate of an atom a against the residual r is
|s|^2 / (|a|^2 |r|^2), s = a^H r (s = Re(a^H r) for signed real coefficients): the share of
the residual energy one atom could remove. Grids of positions are scored, separated peaks
picked, normals chosen from a sphere bank, and the peaks refined by per-candidate
quasi-Newton ascent (damped BFGS, halving backtracks).

certificate_scan (no counterpart in the reference) scores candidates against several scans of
uneven aperture counts in one launch and keeps the best per scan.
"""

from dataclasses import dataclass
from math import ceil

import torch

from .kernels import search

PEAK_POOL = 64  # top scores kept per wanted peak before the separated selection


def certificate(model, params, residual, mode, *, gradient=False, normals=False, scan=True):
    """The squared normalized certificate of every atom and, with gradient, its derivative by
    the centres (and normals). The per-candidate algebra runs in float64 so that the 1e-60
    floor on |a|^2 stays representable (certificate.py:11-31). scan=False keeps the values on
    the correlation kernel that the gradient uses, for comparisons against its values."""
    if gradient:
        corr, norm, d_corr, d_norm = model.correlate_gradient(params, residual, include_normals=normals)
    elif scan:
        corr, norm = model.correlate(params, residual)
    else:
        corr, norm = model.correlate(params, residual, scan=False)
    corr, norm = corr.to(torch.complex128), norm.double().clamp(min=1e-60)
    power = residual.abs().square().sum().double().clamp(min=1e-30)
    s = corr.real if mode == "real" else corr
    value = s.abs().square() / norm / power
    if not gradient:
        return value.float()
    d_corr = d_corr.to(torch.complex128)
    d_norm = torch.where((norm > 1e-60)[:, None], d_norm.double(), 0)
    d_s2 = 2 * s[:, None] * d_corr.real if mode == "real" else 2 * (corr.conj()[:, None] * d_corr).real
    return value.float(), ((d_s2 - value[:, None] * power * d_norm) / norm[:, None] / power).float()


@dataclass(frozen=True)
class CertificateSearchConfig:
    peaks: int = 1024  # separated coarse peaks refined per search
    separation: float = 0.0005  # between coarse peaks, centre units
    iterations: int = 40  # quasi-Newton ascent steps
    trust_radius: float = 1.0  # times each parameter's scale
    backtracks: int = 8


# The one-launch ascent gives one workgroup to each candidate: it fills the chip from this many.
QN_MIN_BATCH = 256


def candidate_grid(lower, upper, spacing, device="cuda"):
    """A regular grid holding both corners of the box; its pitch never exceeds spacing."""
    axes = [torch.linspace(a, b, 1 + ceil((b - a) / spacing), device=device) for a, b in zip(lower, upper, strict=True)]
    return torch.stack(torch.meshgrid(*axes, indexing="ij"), -1).reshape(-1, len(axes))


def select_peaks(scores, centers, count, separation):
    """Greedy best-first indices at least separation apart (a stable sort when separation is
    0). Everything at or above a score level is decided before anything below it, so when
    the PEAK_POOL * count best scores already yield every wanted peak the rest cannot change
    the answer, and the selection kernel runs on that pool (certificate.py:52-70)."""
    if separation == 0:
        ranked = scores.argsort(descending=True, stable=True)[:count]
        return ranked[torch.isfinite(scores[ranked])]
    wanted = min(count, len(scores))
    if len(scores) > PEAK_POOL * wanted:
        level = scores.topk(PEAK_POOL * wanted).values[-1]
        pool = (scores >= level).nonzero().flatten()
        picked = search.peaks(scores[pool], centers[pool], wanted, separation)
        if len(picked) == wanted:
            return pool[picked]
    return search.peaks(scores, centers, count, separation)


def fibonacci_sphere(count, device="cuda"):
    """count nearly uniform unit vectors [count, 3] (golden-angle spiral in z)."""
    k = torch.arange(count, device=device, dtype=torch.float32)
    z = 1 - 2 * (k + 0.5) / count
    turn = k * (torch.pi * (3 - 5**0.5))
    ring = (1 - z.square()).sqrt()
    return torch.stack((ring * turn.cos(), ring * turn.sin(), z), -1)


@torch.no_grad()
def orient_candidates(model, params, residual, count, mode, table=None, chunk=8192, *, batched=False):
    """Give every position the best of its own normal and a count-normal sphere bank: the
    candidates with the chosen normals, and their certificates (certificate.py:73-103).
    batched scores every candidate and picks its normal in one launch, without the
    [C, 1 + count] correlation sums."""
    bank = fibonacci_sphere(count, params["centers"].device)
    power = residual.abs().square().sum().double().clamp(min=1e-30)
    if batched and hasattr(model, "orient_select"):
        value, pick = model.orient_select(params, residual, bank, table, mode == "real")
        normals = torch.where((pick > 0)[:, None], bank[(pick - 1).clamp(min=0)], params["normals"])
        return {**params, "normals": normals}, (value / power).float()
    best_scores, best_normals = [], []
    for first in range(0, len(params["centers"]), chunk):
        part = {key: value[first : first + chunk] for key, value in params.items()}
        corr, norm = model.correlate_orientations(part, residual, bank, table)
        corr, norm = corr.to(torch.complex128), norm.double()
        s = corr.real if mode == "real" else corr
        best, pick = (s.abs().square() / norm.clamp(min=1e-60) / power).max(1)
        normals = part["normals"].clone()
        from_bank = pick > 0
        normals[from_bank] = bank[pick[from_bank] - 1]
        best_scores.append(best.float())
        best_normals.append(normals)
    return {**params, "normals": torch.cat(best_normals)}, torch.cat(best_scores)


class _SearchSpace:
    """The coordinates an ascent moves: centres (and unit normals), their step scales,
    trainable masks and box, as the [4, 6] controls of the trial kernel (scale, mask, lower,
    upper per coordinate)."""

    def __init__(self, candidates, center_spec, normal_spec):
        reference = candidates["centers"]
        self.start = dict(candidates, centers=center_spec.project(reference, reference))
        self.specs = {"centers": center_spec}
        if normal_spec is not None:
            self.specs["normals"] = normal_spec
            self.start["normals"] = normal_spec.project(candidates["normals"], candidates["normals"])
        self.masks = {key: spec.mask(self.start[key]) for key, spec in self.specs.items()}
        controls = reference.new_zeros((4, 6))
        controls[0, :3] = center_spec.scale
        controls[1, :3] = self.masks["centers"]
        controls[2, :3] = reference.new_tensor(-torch.inf if center_spec.lower is None else center_spec.lower)
        controls[3, :3] = reference.new_tensor(torch.inf if center_spec.upper is None else center_spec.upper)
        if normal_spec is not None:
            controls[0, 3:], controls[1, 3:] = normal_spec.scale, 1
        self.controls = controls
        self.with_normals = normal_spec is not None

    def descent(self, params, gradient):
        """The certificate gradient in solver units, tangent and masked, negated (the ascent
        minimizes the negated certificate)."""
        parts = [
            spec.tangent(gradient[:, 3 * i : 3 * i + 3], params[key]) * self.masks[key] * spec.scale
            for i, (key, spec) in enumerate(self.specs.items())
        ]
        return -torch.cat(parts, -1)

    def step(self, params, before):
        return torch.cat([(params[key] - before[key]) / spec.scale for key, spec in self.specs.items()], -1)

    def kernel_spec(self, config, mode, residual):
        """The constants of the one-launch ascent (dsplat_qn_spec of the C ABI)."""
        from ._lib import QnSpec

        ctl = self.controls.cpu()
        centers, normals = self.specs["centers"], self.specs.get("normals")
        spec = QnSpec()
        spec.center_scale = float(centers.scale)
        spec.normal_scale = float(normals.scale) if normals is not None else 1.0
        spec.center_mask[:] = self.masks["centers"].float().cpu().tolist()
        spec.normal_mask[:] = self.masks["normals"].float().cpu().tolist() if normals is not None else [1.0] * 3
        spec.direction_mask[:] = ctl[1].tolist()
        spec.step_scale[:] = ctl[0].tolist()
        spec.lower[:] = ctl[2, :3].tolist()
        spec.upper[:] = ctl[3, :3].tolist()
        spec.width = 6 if normals is not None else 3
        spec.normalize = int(bool(normals.normalize)) if normals is not None else 0
        spec.real_part = int(mode == "real")
        spec.iterations = int(config.iterations)
        spec.backtracks = int(config.backtracks)
        spec.radius = float(config.trust_radius)
        spec.power = float(residual.abs().square().sum().double().clamp(min=1e-30))
        return spec


def _bfgs_update(inverse, step, change, moved, identity):
    """Damped inverse BFGS update per candidate, skipped where the curvature condition fails:
    H <- (I - rho s y^T) H (I - rho y s^T) + rho s s^T, rho = 1 / (s . y)."""
    sy = (step * change).sum(-1)
    usable = moved & (sy > 1e-8 * step.norm(dim=-1) * change.norm(dim=-1)) & (sy > 1e-20)
    rho = torch.where(usable, 1 / sy.clamp(min=1e-30), 0)
    left = identity - rho[:, None, None] * step[:, :, None] * change[:, None, :]
    updated = left @ inverse @ left.transpose(-1, -2) + rho[:, None, None] * step[:, :, None] * step[:, None, :]
    return torch.where(usable[:, None, None], updated, inverse)


def _ascend_on_host(space, params, score, config):
    """The per-candidate quasi-Newton ascent as host-driven batched steps: a direction from the
    inverse Hessian, halving backtracks until each candidate's certificate rises, then the
    BFGS update. Returns the refined params and certificates. Trial certificates come from
    the correlation kernel the gradient's values come from (score(..., scan=False)), so every
    comparison is between values of one kernel. (Scoring all fractions in one launch measured
    slower at the DSFW shapes: most candidates rise at the first fraction.)"""
    value, grad = score(params, True)
    grad = space.descent(params, grad)
    count, width = grad.shape
    identity = torch.eye(width, device=grad.device).expand(count, width, width)
    inverse = identity.clone()
    for _ in range(config.iterations):
        direction = search.direction(inverse, grad, params["normals"], space.controls[1], config.trust_radius)
        before = {key: v.clone() for key, v in params.items()}
        moved = torch.zeros(count, dtype=torch.bool, device=grad.device)
        for halving in range(config.backtracks):
            centers, normals = search.trial(before["centers"], before["normals"], direction, space.controls,
                                            0.5**halving)  # fmt: skip
            trial = dict(before, centers=centers, normals=normals)
            open_ids = torch.where(~moved)[0]
            trial_value = score({key: v[open_ids] for key, v in trial.items()}, scan=False)
            rises = torch.zeros_like(moved)
            rises[open_ids] = torch.isfinite(trial_value) & (trial_value > value[open_ids])
            for key in space.specs:
                params[key][rises] = trial[key][rises]
            moved |= rises
            if moved.all():
                break
        if not moved.any():
            break
        new_value, new_grad = score(params, True)
        new_grad = space.descent(params, new_grad)
        inverse = _bfgs_update(inverse, space.step(params, before), new_grad - grad, moved, identity)
        value, grad = new_value, new_grad
    return params, value


@torch.no_grad()
def search_certificate(
    model, residual, candidates, center_spec, config, mode, *, normal_spec=None,
    initial_scores=None, full_model=None, full_residual=None, batched=False,
):  # fmt: skip
    """Refine separated peaks of the certificate in xyz (and unit normals); the coarse peaks
    stay beside the refined ones and all are ranked, on full_model when given
    (certificate.py:106-201). batched runs every candidate's whole ascent in one launch."""
    space = _SearchSpace(candidates, center_spec, normal_spec)

    def score(params, gradient=False, scan=True):
        return certificate(model, params, residual, mode, gradient=gradient, normals=space.with_normals, scan=scan)

    coarse_scores = score(space.start) if initial_scores is None else initial_scores
    ids = select_peaks(coarse_scores, space.start["centers"], config.peaks, config.separation)
    coarse = {key: value[ids] for key, value in space.start.items()}
    # one workgroup per candidate fills the chip only with enough candidates; a few dozen
    # (the full-aperture refinement) run the host loop, whose launches spread every aperture,
    # unless the model's ascent runs in lockstep rounds whose launches do that too
    if (batched and hasattr(model, "qn_search") and not center_spec.normalize
            and (len(ids) >= QN_MIN_BATCH or getattr(model, "lockstep_search", False))):  # fmt: skip
        centers, normals, value = model.qn_search(coarse, residual, space.kernel_spec(config, mode, residual))
        refined = dict(coarse, centers=centers, normals=normals)
    else:
        refined, value = _ascend_on_host(space, {key: v.clone() for key, v in coarse.items()}, score, config)
    both = {key: torch.cat((coarse[key], refined[key])) for key in refined}
    scores = torch.cat((coarse_scores[ids], value)) if full_model is None else certificate(full_model, both,
                                                                                           full_residual, mode)
    order = scores.argsort(descending=True, stable=True)
    return {key: v[order] for key, v in both.items()}, scores[order]


@torch.no_grad()
def certificate_scan(models, residuals, candidates, mode="complex", top=256, chunk=None):
    """Certificates of every candidate against several scans of one chirp in one launch, and
    the top best per scan (stable: ties keep the lower index), both on the device: the scan,
    a float64 scoring launch and one top selection per scan.

    models are FMCWModels sharing the chirp and material; residuals [A_i, K]. Returns
    (scores [n, C] float32, top ids [n, min(top, C)] int64)."""
    from .kernels.search import certificate_scores, top_descending
    from .kernels.surfel import scan_rows

    config = models[0].scan.fmcw_config
    scans = [(m.scan.scan_positions_m, r) for m, r in zip(models, residuals, strict=True)]
    rows = scan_rows(config, scans, candidates["centers"], candidates["normals"], models[0].fresnel_n, chunk)
    power = torch.stack([r.abs().square().sum().double() for r in residuals]).clamp(min=1e-30)
    scores = certificate_scores(rows, candidates["areas"], power, mode == "real")
    if scores.shape[1] == 0:
        return scores.float(), scores.new_zeros((len(scans), 0), dtype=torch.int64)
    return scores.float(), top_descending(scores, top)
