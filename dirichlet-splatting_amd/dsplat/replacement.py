"""Fixed-budget support replacement.

Public surface of the reference's dsplat/replacement.py:1-270 (utility, replace_one,
regularized_inverse, backward_elimination, pooled_statistics, merge_support).

  utility  the least useful atom gives its slot to the best certificate candidate
  merged   the best candidates join the support in one pooled least-squares system; the
           pooled atoms are then removed one at a time, cheapest loss first, until the budget
           holds again

Either way a support is installed only when its refitted objective is lower. The pooled
normal equations reuse the support's Gram and stream only the new columns; the elimination
runs as one HIP launch per step (no host synchronization) for real coefficients in the
batched form.
"""

from dataclasses import replace

import numpy as np
import torch

from .certificate import certificate
from .kernels import solver as kernels
from .kernels.solver import cholesky_ex
from .parameters import periodic_delta
from .varpro import (
    Fit, _Folded, accept, block_products, fit_coefficients, rendered_loss, row_chunk, screen_bar, shifted_cholesky,
    varpro,
)  # fmt: skip


def _feasible(state, pool, owner):
    """The pool projected onto the parameter constraints; fields it lacks copy slot owner."""
    size = len(pool["centers"])
    feasible = {}
    for key, spec in state.parameters.items():
        reference = state.params[key][owner].expand(size, *state.params[key].shape[1:])
        feasible[key] = spec.project(pool.get(key, reference), reference)
    return feasible


# --- One slot ------------------------------------------------------------------------------


def _inverse_diagonal(gram):
    """(G^-1)_jj, and a mask of the columns that take part in G's null space (None when G
    is positive definite). A singular G uses its pseudoinverse."""
    factor, info = cholesky_ex(gram)
    if int(info) == 0:
        return torch.cholesky_inverse(factor).diag().real, None
    values, vectors = torch.linalg.eigh(gram)
    eps = torch.finfo(values.dtype).eps
    kept = values > values.abs().max() * len(values) * eps
    reciprocal = torch.where(kept, values.clamp(min=torch.finfo(values.dtype).tiny).reciprocal(), 0)
    inverse_diag = (vectors.abs().square() * reciprocal[None]).sum(1)
    in_null = vectors[:, ~kept].abs().square().sum(1) > len(values) * eps * 10
    return inverse_diag, in_null


def utility(gram, coefficients):
    """Per atom: half its relative amplitude plus half its projection efficiency
    1 / sqrt(max(1, (G^-1)_jj G_jj)); a column taking part in the null space of a singular G,
    or an empty one, is fully redundant (replacement.py:46-59)."""
    diagonal = gram.diag().real
    inverse_diag, in_null = _inverse_diagonal(gram)
    efficiency = (1 / (inverse_diag * diagonal).clamp(min=1)).sqrt()
    if in_null is not None:
        efficiency = torch.where(in_null, 0, efficiency)
        efficiency = torch.where(diagonal > 0, efficiency, 0)
    amplitude = coefficients.abs()
    return 0.5 * amplitude / amplitude.max().clamp(min=1e-30) + 0.5 * efficiency


def _exclude_near_support(scores, candidates, state, separation):
    """-inf for candidates closer than separation to an atom of the support (blocks of 256)."""
    support, period = state.params["centers"], state.parameters["centers"].period
    for first in range(0, len(scores), 256):
        gaps = periodic_delta(candidates[first : first + 256, None] - support[None], period).norm(dim=-1)
        scores[first : first + 256][gaps.min(1).values < separation] = -torch.inf


@torch.no_grad()
def replace_one(state, config, pool, screen=None):
    """The best pool candidate tried in the least useful slot: (state, atoms replaced)."""
    owner = int(utility(state.gram, state.coefficients).argmin())
    candidate = _feasible(state, pool, owner)
    scores = certificate(state.model, candidate, state.target - state.prediction, config.coefficients)
    if config.min_separation > 0:
        _exclude_near_support(scores, candidate["centers"], state, config.min_separation)
    scores[~torch.isfinite(scores)] = -torch.inf
    best = int(scores.argmax())
    if not scores[best] > 0:
        return state, 0
    proposal = {key: value.clone() for key, value in state.params.items()}
    for key in proposal:
        proposal[key][owner] = candidate[key][best]
    accepted = accept(state, proposal, config, screen_bar(screen, state.params, config))
    return (state, 0) if accepted is None else (accepted, 1)


# --- Pooled elimination ---------------------------------------------------------------------


def regularized_inverse(h, floor):
    """(h + floor lambda_max(h) I)^-1 in float64, or None (replacement.py:91-101)."""
    factor = shifted_cholesky(h, floor)
    return None if factor is None else torch.cholesky_inverse(factor)


def _eliminate_on_host(inverse, b, count, limited, limit):
    """The elimination as a loop of tensor operations (complex coefficients)."""
    size = len(inverse)
    remaining = inverse.diagonal().real.clone()  # W_jj of the unknowns still in
    alive = torch.ones(size, device=inverse.device, dtype=torch.bool)
    corrections = inverse.new_zeros((size, count))  # the rank-one downdates so far
    removed = torch.zeros(count, device=inverse.device, dtype=torch.long)
    used = torch.zeros((), device=inverse.device, dtype=torch.long)
    tiny = torch.finfo(remaining.dtype).tiny
    for step in range(count):
        loss = b.abs().square() / remaining.clamp(min=tiny)
        barred = ~alive if limited is None else ~alive | (limited & (used >= limit))
        j = loss.masked_fill(barred, torch.inf).argmin()
        column = inverse[:, j] - corrections[:, :step] @ corrections[j, :step].conj()
        pivot = column[j].real.clamp(min=tiny)
        b = b - column * (b[j] / pivot)
        corrections[:, step] = column / pivot.sqrt()
        remaining = remaining - corrections[:, step].abs().square()
        alive[j] = False
        remaining[j] = 1.0
        b[j] = 0
        removed[step] = j
        if limited is not None:
            used = used + limited[j]
    return removed


def backward_elimination(inverse, rhs, count, limited=None, limit=0, *, kernel=False):
    """Remove count unknowns, one at a time, from the least squares whose inverse is W.

    Removing unknown j raises the objective by |b_j|^2 / W_jj with b = W q; each step removes
    the cheapest (ties to the lower index) and keeps the exact inverse of the rest by the
    rank-one downdate W - w w^H / W_jj. At most limit of the unknowns flagged in limited
    leave. Returns the removed indices in order (replacement.py:104-139). With kernel (real
    float64 W) the loop runs on the device, one launch per step."""
    b = inverse @ rhs.to(inverse.dtype)
    if kernel and not inverse.is_complex():
        return kernels.backward_elimination(inverse, b, count, limited, limit)
    return _eliminate_on_host(inverse, b, count, limited, limit)


def _greedy_in_order(clash, free):
    """Walk candidates in score order; one clashing with an earlier kept one is dropped.
    clash [n, n] upper triangular (host), free [n] (host); returns the kept mask."""
    keep = free.copy()
    for i in np.flatnonzero(free):
        if (clash[:i, i] & keep[:i]).any():
            keep[i] = False
    return keep


def _distinct(scores, points, occupied, period, separation, width):
    """Up to width best-scoring candidates that collide neither with the support nor with a
    better kept candidate; a collision is a distance below separation, or exactly zero (an atom
    pooled twice makes the system singular for nothing). Candidates are taken in blocks of
    1024 in score order (replacement.py:142-175)."""

    def gaps(a, b):
        return periodic_delta(a[:, None] - b[None], period).norm(dim=-1)

    def apart(distance):
        return (distance >= separation) & (distance > 0)

    ranked = scores.argsort(descending=True, stable=True)
    ranked = ranked[torch.isfinite(scores[ranked])]
    chosen, total = [], 0
    for first in range(0, len(ranked), 1024):
        if total >= width:
            break
        ids = ranked[first : first + 1024]
        block = points[ids]
        free = apart(gaps(block, occupied).min(1).values)
        for earlier in chosen:
            free &= apart(gaps(block, points[earlier]).min(1).values)
        within = gaps(block, block)
        clash = ((within < separation) | (within == 0)).triu(1)
        if separation <= 0:
            # collisions are exact coincidences, an equivalence relation: the in-order walk
            # keeps the first free member of each class, which one masked reduction decides
            free &= ~(clash & free[:, None]).any(0)
        elif bool(clash[free][:, free].any()):
            free = torch.as_tensor(_greedy_in_order(clash.cpu().numpy(), free.cpu().numpy()), device=free.device)
        chosen.append(ids[free])
        total += int(free.sum())
    return torch.cat(chosen)[:width] if chosen else ranked[:0]


def pooled_statistics(state, pool, config):
    """Gram and right-hand side over the support followed by the new pooled atoms; only the
    new columns are streamed, the support block is the state's (replacement.py:178-199)."""
    model, count = state.model, len(state.coefficients)
    total, added = len(pool["centers"]), len(pool["centers"]) - count
    gram = state.gram.new_empty((total, total))
    rhs = state.rhs.new_empty(total)
    gram[:count, :count], rhs[:count] = state.gram, state.rhs
    if config.real and getattr(config, "batched", False) and hasattr(model, "adjoint_correlation"):
        # cross products of the new columns with all of them on the MFMA block products, their
        # right-hand side from one correlation
        acc = _Folded(total, added, state.target.device)
        block_products(model, pool, state.target, config, acc, col0=count, known=state.params)
        cross = (acc.result() / model.size).float()
        fresh = {key: value[count:] for key, value in pool.items()}
        right = (model.adjoint_correlation(fresh, state.target).real.double() / model.size).float()
    else:
        cross = state.gram.new_zeros((total, added))
        right = state.rhs.new_zeros(added)
        for a, y in model.blocks(pool, state.target, config.real, row_chunk(model, total, config)):
            cross.addmm_(a.H, a[:, count:])
            right.addmv_(a[:, count:].H, y)
        cross, right = cross / model.size, right / model.size
    gram[:, count:], gram[count:, :] = cross, cross.H
    gram[count:, count:] = (cross[count:] + cross[count:].H) * 0.5
    rhs[count:] = right
    return gram, rhs


def _settle(state, config, pool, gram, rhs, ids, owners):
    """The refit of the support pool[ids] (slots owners took pooled atoms), or None. Fields
    whose constraints a pooled atom's slot changes are projected, and then the geometry is
    refitted from scratch; otherwise the pooled system's rows serve."""
    trial = {key: value[ids] for key, value in pool.items()}
    moved = False
    for key, spec in state.parameters.items():
        if not spec.mask(trial[key]).all():
            fixed = spec.project(trial[key][owners], state.params[key][owners])
            moved = moved or not torch.equal(fixed, trial[key][owners])
            trial[key][owners] = fixed
    if moved:
        fit = varpro(state.model, trial, state.target, config)
    else:
        sub_gram, sub_rhs = gram[ids[:, None], ids[None, :]], rhs[ids]
        b = fit_coefficients(sub_gram, sub_rhs, config)
        fit = Fit(b, *rendered_loss(state.model, trial, b, state.target, config.ridge), sub_gram, sub_rhs)
    return trial, fit


@torch.no_grad()
def merge_support(state, config, pool, scores=None):
    """Pool the best candidates with the support, fit once, eliminate back to the budget; a
    rejected support is retried with half as many old atoms allowed to leave
    (replacement.py:202-270). Returns the state and the number of atoms replaced."""
    support = state.params["centers"]
    count = len(support)
    candidate = _feasible(state, pool, 0)
    if scores is None:
        scores = certificate(state.model, candidate, state.target - state.prediction, config.coefficients)
    joining = _distinct(scores, candidate["centers"], support, state.parameters["centers"].period,
                        config.min_separation, config.batch)  # fmt: skip
    if not len(joining):
        return state, 0
    pooled = {key: torch.cat((value, candidate[key][joining])) for key, value in state.params.items()}
    gram, rhs = pooled_statistics(state, pooled, config)
    h = (gram.real if config.real else gram) + config.ridge * torch.eye(len(gram), dtype=gram.dtype,
                                                                          device=gram.device)  # fmt: skip
    inverse = regularized_inverse(h, len(h) * torch.finfo(h.real.dtype).eps)
    if inverse is None:
        return state, 0
    q = rhs.real if config.real else rhs
    original = torch.arange(len(q), device=q.device) < count
    on_device = getattr(config, "batched", False)
    limit = None  # how many original atoms may leave (None: no bound)
    while True:
        removed = backward_elimination(inverse, q, len(joining), None if limit is None else original, limit or 0,
                                       kernel=on_device)  # fmt: skip
        gone = torch.zeros(len(q), device=q.device, dtype=torch.bool)
        gone[removed] = True
        owners = gone[:count].nonzero().flatten()
        if not len(owners):
            return state, 0
        # each vacated slot takes the next surviving pooled atom, in pool order
        ids = torch.arange(count, device=q.device)
        ids[owners] = (~gone[count:]).nonzero().flatten() + count
        trial, fit = _settle(state, config, pooled, gram, rhs, ids, owners)
        if fit.objective < state.objective:
            cache = getattr(state.model, "split_cache", None)
            if cache is not None:
                cache.adopt(pooled, trial, owners, ids[owners])
            return replace(state, params=trial, **fit._asdict()), len(owners)
        limit = len(owners) // 2
        if limit < 1:
            return state, 0
