"""Spatial candidate search: hierarchical 3-D grids and continuous certificate refinement.

Same interface as dsplat/spatial.py of the reference (spatial.py:26-178). Each step
builds a fresh candidate set inside the centers' box: a range-gated coarse grid plus
cached peaks, finer cubes around separated peaks, quasi-Newton refinement on an
aperture subset, then on every aperture.
"""

from dataclasses import dataclass, field, replace
from math import ceil, prod

import torch

from .certificate import (
    CertificateSearchConfig,
    candidate_grid,
    certificate,
    orient_candidates,
    search_certificate,
    select_peaks,
)


@dataclass(frozen=True)
class SpatialSearchConfig:
    coarse_spacing: float = 0.002
    levels: tuple[tuple[float, float], ...] = ((0.002, 0.0005), (0.0005, 0.000125))
    proposal_apertures: int = 384
    refine_apertures: int = 768
    max_candidates: int = 2_000_000
    range_power: float = 0.03
    range_margin_cells: float = 1.5
    range_fraction: float = 0.7
    ungated_after: int = 2
    normal_candidates: int = 24
    joint_normals: bool = True
    fft_oversample: int = 8
    cache_peaks: int = 256
    full_peaks: int = 64
    full_iterations: int = 8
    continuous: CertificateSearchConfig = field(default_factory=CertificateSearchConfig)


def _grid_size(lower, upper, pitch):
    return prod(1 + ceil(float(b - a) / pitch) for a, b in zip(lower, upper, strict=True))


@torch.no_grad()
def propose(state, config, search):
    """Freshly refined candidates, their scores best first, and the peaks to cache.

    After failed steps the coarse grid gets finer and shifts, more peaks are kept and the
    range gate loosens (spatial.py:47-178).
    """
    centers = state.params["centers"]
    spec, normal_spec = state.parameters["centers"], state.parameters["normals"]
    free_normals = bool(normal_spec.mask(state.params["normals"]).any())
    refine_normals = normal_spec if free_normals and search.joint_normals else None
    failures, mode = state.search_failures, config.coefficients
    residual = state.target - state.prediction
    ids, model = state.model.sample_apertures(search.proposal_apertures, config.seed + state.iteration)
    r = residual[ids]
    banked = free_normals and search.normal_candidates
    table = model.range_table(r, search.fft_oversample) if banked and search.fft_oversample else None

    def score(candidates):
        if banked:
            oriented, scores = orient_candidates(
                model, candidates, r, search.normal_candidates, mode, table, batched=getattr(config, "batched", False)
            )
            candidates["normals"] = oriented["normals"]
            return scores
        return certificate(model, candidates, r, mode)

    def atoms(points, normals=None):
        result = {
            key: value[:1].expand(len(points), *value.shape[1:]).clone() for key, value in state.params.items()
        }
        result["centers"] = spec.project(points, result["centers"])
        if normals is not None:
            result["normals"] = normal_spec.project(normals, result["normals"])
        return result

    peaks = min(search.continuous.peaks * (1 + failures), 4 * search.continuous.peaks)
    separation = search.continuous.separation
    lower = centers.new_tensor(spec.lower).expand(3).clone()
    upper = centers.new_tensor(spec.upper).expand(3).clone()
    frozen = ~spec.mask(centers)
    lower[frozen] = upper[frozen] = centers[0, frozen]

    spacing = search.coarse_spacing / (1 + min(failures, 2))
    while _grid_size(lower, upper, spacing) > search.max_candidates:
        spacing *= 1.1
    points = candidate_grid(lower.tolist(), upper.tolist(), spacing)
    if failures:
        shift = points.new_tensor([0.37, 0.61, 0.23]) * spacing * ((failures % 3) + 1) / 3
        points = spec.project(points + shift, points)
    if failures < search.ungated_after:
        keep = model.range_gate(
            points, state.target[ids], relative_power=search.range_power / (1 + failures),
            margin_cells=search.range_margin_cells * (1 + failures), min_fraction=search.range_fraction,
        )  # fmt: skip
        if keep.any():
            points = points[keep]
    cached = 0
    if state.peak_cache is not None:
        cached = min(len(state.peak_cache["centers"]), search.max_candidates - len(points))
        points = torch.cat((points, state.peak_cache["centers"][:cached]))
    candidates = atoms(points)
    if cached:
        candidates["normals"][-cached:] = normal_spec.project(
            state.peak_cache["normals"][:cached], candidates["normals"][-cached:]
        )

    scores = score(candidates)
    selected = select_peaks(scores, candidates["centers"], peaks, separation)
    candidates = {key: value[selected] for key, value in candidates.items()}
    for radius, pitch in search.levels:
        lo = [-radius if active else 0 for active in ~frozen]
        hi = [radius if active else 0 for active in ~frozen]
        while _grid_size(lo, hi, pitch) > search.max_candidates:
            pitch *= 1.1
        offsets = candidate_grid(lo, hi, pitch)
        budget = max(1, search.max_candidates // len(offsets))
        candidates = {key: value[:budget] for key, value in candidates.items()}
        points = (candidates["centers"][:, None] + offsets[None]).reshape(-1, 3)
        candidates = atoms(points, candidates["normals"].repeat_interleave(len(offsets), dim=0))
        scores = score(candidates)
        selected = select_peaks(scores, candidates["centers"], peaks, min(separation, radius / 2))
        candidates = {key: value[selected] for key, value in candidates.items()}
    initial = scores[selected] if table is None else None
    if search.refine_apertures > search.proposal_apertures:
        refine_ids, model = state.model.sample_apertures(
            search.refine_apertures, config.seed + state.iteration + 1000003
        )
        r, initial = residual[refine_ids], None
    params, scores = search_certificate(
        model, r, candidates, spec, replace(search.continuous, peaks=peaks, separation=0), mode,
        normal_spec=refine_normals, initial_scores=initial, full_model=state.model, full_residual=residual,
        batched=getattr(config, "batched", False),
    )  # fmt: skip
    if search.full_iterations:
        full = select_peaks(scores, params["centers"], search.full_peaks, min(separation, spec.scale / 2))
        refined, refined_scores = search_certificate(
            state.model, residual, {key: value[full] for key, value in params.items()}, spec,
            replace(search.continuous, peaks=search.full_peaks, iterations=search.full_iterations, separation=0),
            mode, normal_spec=refine_normals, initial_scores=scores[full], batched=getattr(config, "batched", False),
        )  # fmt: skip
        scores = torch.cat((refined_scores, scores))
        order = scores.argsort(descending=True, stable=True)
        params = {key: torch.cat((value, params[key]))[order] for key, value in refined.items()}
        scores = scores[order]
    if not search.cache_peaks:
        return params, scores, None
    cache = select_peaks(scores, params["centers"], search.cache_peaks, separation)
    return params, scores, {key: value[cache].clone() for key, value in params.items()}
