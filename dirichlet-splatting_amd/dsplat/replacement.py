"""Fixed-budget support replacement (replacement.py of the reference).

utility: replace the least useful atom by the best certificate candidate.
merged: pool candidates with the support, rank every atom by the cost of deleting it
from one pooled fit, keep the budget's cheapest-to-lose complement. Both install a
support only if its refitted objective is lower.
"""

from dataclasses import replace

import torch

from .certificate import certificate
from .kernels import solver as elimination
from .parameters import periodic_delta
from .kernels.blas import gemm_tn
from .kernels.solver import cholesky_ex
from .varpro import (
    FOLD, SYRK_MODE, Fit, accept, fit_coefficients, rendered_loss, row_chunk, screen_bar, shifted_cholesky, varpro,
)


def _feasible(state, pool, owner):
    """The pool as feasible atoms; fields the pool lacks copy slot `owner`."""
    count = len(pool["centers"])
    out = {}
    for key, spec in state.parameters.items():
        reference = state.params[key][owner].expand(count, *state.params[key].shape[1:])
        out[key] = spec.project(pool.get(key, reference), reference)
    return out


def utility(gram, coefficients):
    """Half relative amplitude, half unregularized projection efficiency per atom.

    Efficiency 1 / sqrt(max(1, (G^-1)_jj G_jj)); for a singular Gram a column is
    redundant exactly when its coordinate takes part in the null space, and the other
    columns use the pseudoinverse (replacement.py:46-59).
    """
    factor, info = cholesky_ex(gram)
    diag = gram.diag().real
    if int(info) == 0:
        inverse_diag = torch.cholesky_inverse(factor).diag().real
        efficiency = (1 / (inverse_diag * diag).clamp(min=1)).sqrt()
    else:
        values, vectors = torch.linalg.eigh(gram)
        eps = torch.finfo(values.dtype).eps
        active = values > values.abs().max() * len(values) * eps
        inv = torch.where(active, values.clamp(min=torch.finfo(values.dtype).tiny).reciprocal(), 0)
        inverse_diag = (vectors.abs().square() * inv[None]).sum(1)
        null_weight = vectors[:, ~active].abs().square().sum(1)
        efficiency = (1 / (inverse_diag * diag).clamp(min=1)).sqrt()
        efficiency = torch.where(null_weight > len(values) * eps * 10, 0, efficiency)
        efficiency = torch.where(diag > 0, efficiency, 0)
    magnitude = coefficients.abs()
    return 0.5 * magnitude / magnitude.max().clamp(min=1e-30) + 0.5 * efficiency


@torch.no_grad()
def replace_one(state, config, pool, screen=None):
    """Try the best pool candidate in the least useful slot; the state and atoms replaced."""
    owner = int(utility(state.gram, state.coefficients).argmin())
    candidate = _feasible(state, pool, owner)
    residual = state.target - state.prediction
    scores = certificate(state.model, candidate, residual, config.coefficients)
    if config.min_separation > 0:
        centers, period = state.params["centers"], state.parameters["centers"].period
        for start in range(0, len(scores), 256):
            delta = periodic_delta(candidate["centers"][start : start + 256, None] - centers[None], period)
            near = delta.norm(dim=-1).min(1).values < config.min_separation
            scores[start : start + 256][near] = -torch.inf
    scores[~torch.isfinite(scores)] = -torch.inf
    index = int(scores.argmax())
    if not scores[index] > 0:
        return state, 0
    proposal = {key: value.clone() for key, value in state.params.items()}
    for key in proposal:
        proposal[key][owner] = candidate[key][index]
    accepted = accept(state, proposal, config, screen_bar(screen, state.params, config))
    return (state, 0) if accepted is None else (accepted, 1)


def regularized_inverse(h, floor):
    """Inverse of h + floor lambda_max(h) I in float64, or None (replacement.py:91-101)."""
    factor = shifted_cholesky(h, floor)
    return None if factor is None else torch.cholesky_inverse(factor)


def backward_elimination(inverse, rhs, count, limited=None, limit=0, *, kernel=False):
    """Remove `count` unknowns one at a time from the least squares with inverse W.

    Deleting unknown j raises the objective by |b_j|^2 / W_jj with b = W q; each step
    deletes the cheapest and keeps the exact inverse of the rest as rank-one
    corrections W - w w^H / W_jj. At most `limit` of the unknowns marked in `limited`
    leave. Returns the removed indices in order (replacement.py:104-139). With kernel
    (real float64 W) the whole loop runs as one HIP launch instead of a host loop.
    """
    size = len(inverse)
    b = inverse @ rhs.to(inverse.dtype)
    if kernel and not inverse.is_complex():
        return elimination.backward_elimination(inverse, b, count, limited, limit)
    diagonal = inverse.diagonal().real.clone()
    alive = torch.ones(size, device=inverse.device, dtype=torch.bool)
    updates = inverse.new_zeros((size, count))
    removed = torch.zeros(count, device=inverse.device, dtype=torch.long)
    spent = torch.zeros((), device=inverse.device, dtype=torch.long)
    tiny = torch.finfo(diagonal.dtype).tiny
    for step in range(count):
        cost = b.abs().square() / diagonal.clamp(min=tiny)
        blocked = ~alive if limited is None else ~alive | (limited & (spent >= limit))
        index = cost.masked_fill(blocked, torch.inf).argmin()
        column = inverse[:, index] - updates[:, :step] @ updates[index, :step].conj()
        pivot = column[index].real.clamp(min=tiny)
        b = b - column * (b[index] / pivot)
        scaled = column / pivot.sqrt()
        updates[:, step] = scaled
        diagonal = diagonal - scaled.abs().square()
        alive[index] = False
        diagonal[index] = 1.0
        b[index] = 0
        removed[step] = index
        if limited is not None:
            spent = spent + limited[index]
    return removed


def _distinct(scores, points, occupied, period, separation, width):
    """Best-scoring candidates colliding neither with the support nor with each other;
    a collision is a distance below separation or exactly zero (replacement.py:142-175)."""

    def distance(a, b):
        return periodic_delta(a[:, None] - b[None], period).norm(dim=-1)

    order = scores.argsort(descending=True, stable=True)
    order = order[torch.isfinite(scores[order])]
    kept, found = [], 0
    for start in range(0, len(order), 1024):
        if found >= width:
            break
        ids = order[start : start + 1024]
        block = points[ids]
        near = distance(block, occupied).min(1).values
        free = (near >= separation) & (near > 0)
        for previous in kept:
            near = distance(block, points[previous]).min(1).values
            free &= (near >= separation) & (near > 0)
        inside = distance(block, block)
        clash = ((inside < separation) | (inside == 0)).triu(1)
        if separation <= 0:
            # collisions are exact coincidences, an equivalence: the in-order greedy keeps
            # the first free member of each class, so one masked reduction decides it
            free &= ~(clash & free[:, None]).any(0)
        elif bool(clash[free][:, free].any()):
            for i in free.nonzero().flatten().tolist():
                if bool((clash[:i, i] & free[:i]).any()):
                    free[i] = False
        kept.append(ids[free])
        found += int(free.sum())
    return torch.cat(kept)[:width] if kept else order[:0]


def _pooled_cross(model, pool, count, target, config):
    """Real cross products of the new pooled columns with all of them on the split SYRK
    tiles (folded into float64 like the Gram), and their rhs from one correlation."""
    total = len(pool["centers"])
    mode = SYRK_MODE[config.gram]
    part = torch.zeros((total, total - count), dtype=torch.float32, device=target.device)
    folded = torch.zeros_like(part, dtype=torch.float64)
    chunk = row_chunk(model, total, config)
    fused = model.projected_blocks(pool, chunk, column_max=True) if hasattr(model, "projected_blocks") else None
    stream = fused if fused is not None else ((a, None) for a, _ in model.blocks(pool, target, True, chunk))
    for index, (a, maxima) in enumerate(stream):
        gemm_tn(a.contiguous(), count, part, mode, maxima)
        if (index + 1) % FOLD == 0:
            folded += part
            part.zero_()
    folded += part
    fresh = {key: value[count:] for key, value in pool.items()}
    right = model.correlate(fresh, target)[0].real.double()
    return (folded / model.size).float(), (right / model.size).float()


def pooled_statistics(state, pool, config):
    """Gram and rhs over the support followed by the pooled new atoms; only the new
    columns are streamed, the support block is the state's (replacement.py:178-199)."""
    model, count = state.model, len(state.coefficients)
    total, added = len(pool["centers"]), len(pool["centers"]) - count
    gram = state.gram.new_empty((total, total))
    rhs = state.rhs.new_empty(total)
    gram[:count, :count], rhs[:count] = state.gram, state.rhs
    if config.real and getattr(config, "batched", False) and hasattr(model, "correlate"):
        cross, right = _pooled_cross(model, pool, count, state.target, config)
    else:
        cross = state.gram.new_zeros((total, added))
        right = state.rhs.new_zeros(added)
        for a, y in model.blocks(pool, state.target, config.real, row_chunk(model, total, config)):
            new = a[:, count:]
            cross.addmm_(a.H, new)
            right.addmv_(new.H, y)
        cross, right = cross / model.size, right / model.size
    gram[:, count:], gram[count:, :] = cross, cross.H
    gram[count:, count:] = (cross[count:] + cross[count:].H) * 0.5
    rhs[count:] = right
    return gram, rhs


@torch.no_grad()
def merge_support(state, config, pool, scores=None):
    """Pool the best candidates with the support, fit once, prune back to the budget; a
    rejected support is retried with half as many old atoms allowed to leave
    (replacement.py:202-270). Returns the state and the number of atoms replaced."""
    centers = state.params["centers"]
    count = len(centers)
    candidate = _feasible(state, pool, 0)
    if scores is None:
        residual = state.target - state.prediction
        scores = certificate(state.model, candidate, residual, config.coefficients)
    chosen = _distinct(
        scores, candidate["centers"], centers, state.parameters["centers"].period, config.min_separation, config.batch
    )
    if not len(chosen):
        return state, 0
    pool = {key: torch.cat((value, candidate[key][chosen])) for key, value in state.params.items()}
    gram, rhs = pooled_statistics(state, pool, config)
    real = config.real
    h = (gram.real if real else gram) + config.ridge * torch.eye(len(gram), dtype=gram.dtype, device=gram.device)
    inverse = regularized_inverse(h, len(h) * torch.finfo(h.real.dtype).eps)
    if inverse is None:
        return state, 0
    q = rhs.real if real else rhs
    old = torch.arange(len(q), device=q.device) < count
    limit = None
    while True:
        removed = backward_elimination(
            inverse, q, len(chosen), None if limit is None else old, limit or 0, kernel=getattr(config, "batched", False)
        )
        gone = torch.zeros(len(q), device=q.device, dtype=torch.bool)
        gone[removed] = True
        owners = gone[:count].nonzero().flatten()
        if not len(owners):
            return state, 0
        ids = torch.arange(count, device=q.device)
        ids[owners] = (~gone[count:]).nonzero().flatten() + count
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
            trial_gram, trial_rhs = gram[ids[:, None], ids[None, :]], rhs[ids]
            b = fit_coefficients(trial_gram, trial_rhs, config)
            fit = Fit(b, *rendered_loss(state.model, trial, b, state.target, config.ridge), trial_gram, trial_rhs)
        if fit.objective < state.objective:
            return replace(state, params=trial, **fit._asdict()), len(owners)
        limit = len(owners) // 2
        if limit < 1:
            return state, 0
