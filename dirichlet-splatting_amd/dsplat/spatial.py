"""Spatial candidate search: hierarchical 3-D grids and continuous certificate refinement.

Public surface of the reference's dsplat/spatial.py:26-178 (SpatialSearchConfig, propose).
Every step proposes a fresh candidate set inside the centres' box, in stages:

  1. a range-gated coarse grid plus the peaks cached by the previous step, scored on an
     aperture subset (with the best normal of a sphere bank where normals are free);
  2. for each refinement level, a finer cube of offsets around the separated peaks;
  3. quasi-Newton refinement of the peaks on a larger aperture subset, ranked on every
     aperture;
  4. a short quasi-Newton refinement of the best few on every aperture.

Failed steps make the next coarse grid finer and shifted, keep more peaks and loosen the
range gate.
"""

from dataclasses import dataclass, field, replace
from math import ceil, prod

import torch

from .certificate import CertificateSearchConfig, candidate_grid, certificate, orient_candidates, search_certificate, \
    select_peaks  # fmt: skip


@dataclass(frozen=True)
class SpatialSearchConfig:
    coarse_spacing: float = 0.002  # coarse grid pitch, metres
    levels: tuple[tuple[float, float], ...] = ((0.002, 0.0005), (0.0005, 0.000125))  # (cube radius, pitch)
    proposal_apertures: int = 384  # aperture subset that scores the grids
    refine_apertures: int = 768  # aperture subset of the quasi-Newton refinement
    max_candidates: int = 2_000_000
    range_power: float = 0.03  # range gate: power relative to each profile's peak
    range_margin_cells: float = 1.5  # range gate dilation, range resolutions
    range_fraction: float = 0.7  # range gate: share of apertures a point must hit
    ungated_after: int = 2  # failed steps after which the gate is dropped
    normal_candidates: int = 24  # sphere bank of normals per position
    joint_normals: bool = True  # refine unit normals together with the centres
    fft_oversample: int = 8  # FFT range table positions per bin (0: exact sums)
    cache_peaks: int = 256  # peaks carried to the next step
    full_peaks: int = 64  # peaks refined on every aperture
    full_iterations: int = 8
    continuous: CertificateSearchConfig = field(default_factory=CertificateSearchConfig)


def _grid_size(lower, upper, pitch):
    return prod(1 + ceil(float(b - a) / pitch) for a, b in zip(lower, upper, strict=True))


def _grid_within(lower, upper, pitch, budget):
    """candidate_grid of the box, the pitch grown by 10% steps until it has at most budget points."""
    while _grid_size(lower, upper, pitch) > budget:
        pitch *= 1.1
    return candidate_grid(lower, upper, pitch), pitch


def _take(candidates, ids):
    return {key: value[ids] for key, value in candidates.items()}


class _Proposal:
    """The context of one step's search: the state, the aperture subset and its residual,
    the box, and how candidates are scored."""

    def __init__(self, state, config, search):
        self.state, self.config, self.search = state, config, search
        self.center_spec, self.normal_spec = state.parameters["centers"], state.parameters["normals"]
        free_normals = bool(self.normal_spec.mask(state.params["normals"]).any())
        self.refine_normals = self.normal_spec if free_normals and search.joint_normals else None
        self.failures, self.mode = state.search_failures, config.coefficients
        self.batched = getattr(config, "batched", False)
        self.residual = state.target - state.prediction
        self.ids, self.model = state.model.sample_apertures(search.proposal_apertures, config.seed + state.iteration)
        self.subset_residual = self.residual[self.ids]
        self.banked = free_normals and search.normal_candidates
        self.table = (self.model.range_table(self.subset_residual, search.fft_oversample)
                      if self.banked and search.fft_oversample else None)  # fmt: skip
        self.peaks = min(search.continuous.peaks * (1 + self.failures), 4 * search.continuous.peaks)
        self.separation = search.continuous.separation
        centers = state.params["centers"]
        self.lower = centers.new_tensor(self.center_spec.lower).expand(3).clone()
        self.upper = centers.new_tensor(self.center_spec.upper).expand(3).clone()
        self.frozen = ~self.center_spec.mask(centers)
        self.lower[self.frozen] = self.upper[self.frozen] = centers[0, self.frozen]

    def score(self, candidates):
        """Certificates on the subset; with a normal bank the best normal replaces each
        candidate's (in place)."""
        if self.banked:
            oriented, scores = orient_candidates(self.model, candidates, self.subset_residual,
                                                 self.search.normal_candidates, self.mode, self.table,
                                                 batched=self.batched)  # fmt: skip
            candidates["normals"] = oriented["normals"]
            return scores
        return certificate(self.model, candidates, self.subset_residual, self.mode)

    def atoms(self, points, normals=None):
        """Feasible candidates at points; fields other than the centre (and normal) copy atom 0."""
        out = {key: value[:1].expand(len(points), *value.shape[1:]).clone() for key, value in self.state.params.items()}
        out["centers"] = self.center_spec.project(points, out["centers"])
        if normals is not None:
            out["normals"] = self.normal_spec.project(normals, out["normals"])
        return out

    def coarse(self):
        """Stage 1: the gated coarse grid and the cached peaks, scored and thinned to peaks."""
        search, failures = self.search, self.failures
        points, spacing = _grid_within(self.lower.tolist(), self.upper.tolist(),
                                       search.coarse_spacing / (1 + min(failures, 2)), search.max_candidates)
        if failures:
            shift = points.new_tensor([0.37, 0.61, 0.23]) * spacing * ((failures % 3) + 1) / 3
            points = self.center_spec.project(points + shift, points)
        if failures < search.ungated_after:
            gate = self.model.range_gate(
                points, self.state.target[self.ids], relative_power=search.range_power / (1 + failures),
                margin_cells=search.range_margin_cells * (1 + failures), min_fraction=search.range_fraction,
            )  # fmt: skip
            if gate.any():
                points = points[gate]
        cache = self.state.peak_cache
        carried = 0 if cache is None else min(len(cache["centers"]), search.max_candidates - len(points))
        if carried:
            points = torch.cat((points, cache["centers"][:carried]))
        candidates = self.atoms(points)
        if carried:
            candidates["normals"][-carried:] = self.normal_spec.project(cache["normals"][:carried],
                                                                        candidates["normals"][-carried:])
        scores = self.score(candidates)
        picked = select_peaks(scores, candidates["centers"], self.peaks, self.separation)
        return _take(candidates, picked), scores[picked]

    def refine_levels(self, candidates, scores):
        """Stage 2: a cube of offsets around every peak per level, rescored and thinned."""
        search = self.search
        for radius, pitch in search.levels:
            low = [-radius if moving else 0 for moving in ~self.frozen]
            high = [radius if moving else 0 for moving in ~self.frozen]
            offsets, _ = _grid_within(low, high, pitch, search.max_candidates)
            candidates = _take(candidates, slice(0, max(1, search.max_candidates // len(offsets))))
            points = (candidates["centers"][:, None] + offsets[None]).reshape(-1, 3)
            candidates = self.atoms(points, candidates["normals"].repeat_interleave(len(offsets), dim=0))
            level_scores = self.score(candidates)
            picked = select_peaks(level_scores, candidates["centers"], self.peaks, min(self.separation, radius / 2))
            candidates, scores = _take(candidates, picked), level_scores[picked]
        return candidates, scores

    def continuous(self, candidates, scores):
        """Stage 3: quasi-Newton refinement on the refinement subset, ranked on every aperture."""
        search, state = self.search, self.state
        model, residual = self.model, self.subset_residual
        initial = scores if self.table is None else None
        if search.refine_apertures > search.proposal_apertures:
            refine_ids, model = state.model.sample_apertures(search.refine_apertures,
                                                             self.config.seed + state.iteration + 1000003)
            residual, initial = self.residual[refine_ids], None
        return search_certificate(
            model, residual, candidates, self.center_spec, replace(search.continuous, peaks=self.peaks, separation=0),
            self.mode, normal_spec=self.refine_normals, initial_scores=initial, full_model=state.model,
            full_residual=self.residual, batched=self.batched,
        )  # fmt: skip

    def on_every_aperture(self, params, scores):
        """Stage 4: the best full_peaks refined on every aperture, merged with the rest."""
        search = self.search
        best = select_peaks(scores, params["centers"], search.full_peaks, min(self.separation, self.center_spec.scale / 2))
        refined, refined_scores = search_certificate(
            self.state.model, self.residual, _take(params, best), self.center_spec,
            replace(search.continuous, peaks=search.full_peaks, iterations=search.full_iterations, separation=0),
            self.mode, normal_spec=self.refine_normals, initial_scores=scores[best], batched=self.batched,
        )  # fmt: skip
        merged = torch.cat((refined_scores, scores))
        order = merged.argsort(descending=True, stable=True)
        return {key: torch.cat((value, params[key]))[order] for key, value in refined.items()}, merged[order]


@torch.no_grad()
def propose(state, config, search):
    """Freshly refined candidates, their scores best first, and the peaks to cache
    (spatial.py:47-178)."""
    step = _Proposal(state, config, search)
    candidates, scores = step.refine_levels(*step.coarse())
    params, scores = step.continuous(candidates, scores)
    if search.full_iterations:
        params, scores = step.on_every_aperture(params, scores)
    if not search.cache_peaks:
        return params, scores, None
    cached = select_peaks(scores, params["centers"], search.cache_peaks, step.separation)
    return params, scores, {key: value[cached].clone() for key, value in params.items()}
