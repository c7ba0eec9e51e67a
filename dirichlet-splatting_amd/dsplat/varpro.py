"""Variable projection: streamed normal equations, coefficient elimination and acceptance.

Objective mean(|A b - y|^2) + ridge |b|^2 (varpro.py:1-6 of the reference). Complex
coefficients use the full Gram G = A^H A / M, signed-real ones Re(G) and Re(q). A new
geometry is accepted only if its refitted objective is lower.

Gram backends (DSFWConfig.gram) of the real-coefficient normal equations: the row
stream of the reference (row-projected design blocks, a SYRK per block, folded into an
FP64 total every FOLD blocks; varpro.py:82-120) with the SYRK on
  "plain"   FP32 MFMA operands (the reference's cuBLAS SSYRK)
  "bf16x3"  BF16 hi + lo split operands, three MFMA products, FP32 accumulation
  "fp16x3"  FP16 hi + lo split operands with per-column scaling, likewise
"""

from dataclasses import replace
from typing import NamedTuple

import torch

from .kernels.blas import syrk
from .kernels.reduce import energy
from .kernels.solver import cholesky_ex


class Fit(NamedTuple):
    """VarPro solution of one geometry; the field names match DSFWState."""

    coefficients: torch.Tensor
    prediction: torch.Tensor
    objective: float
    gram: torch.Tensor  # Re(A^H A) / M for real coefficients, A^H A / M otherwise
    rhs: torch.Tensor


POWER_STEPS = 24
SHIFT_RETRIES = 4
FOLD = 8  # float32 SYRK blocks per float64 fold of the Gram


def largest_eigenvalue(h):
    """Power-iteration estimate of lambda_max of a PSD h from a fixed, reproducible start."""
    golden = 0.6180339887
    v = ((torch.arange(len(h), device=h.device, dtype=torch.float64) * golden) % 1 + 0.5).to(h.dtype)
    top = v.norm()
    for _ in range(POWER_STEPS):
        v = h @ v
        top = v.norm()
        v = v / top.clamp(min=torch.finfo(top.dtype).tiny)
    return top


def shifted_cholesky(h, floor):
    """FP64 Cholesky factor of h + shift I, shift = floor lambda_max(h), growing 1, 10, 91...
    times while factorization fails; None after SHIFT_RETRIES failures (varpro.py:44-59)."""
    matrix = h.to(torch.complex128 if h.is_complex() else torch.float64, copy=True)
    shift = floor * largest_eigenvalue(matrix)
    for _ in range(SHIFT_RETRIES):
        matrix.diagonal().add_(shift)
        factor, info = cholesky_ex(matrix)
        if int(info) == 0:
            return factor
        shift = shift * 9
    return None


def solve(h, rhs):
    """Hermitian solve; a singular system is shifted to the level where the pseudoinverse
    truncates, len(h) eps lambda_max, which moves x only along near-null directions
    (varpro.py:62-79). An all-zero system gives zeros."""
    q = rhs[:, None] if rhs.ndim == 1 else rhs
    factor, info = cholesky_ex(h)
    if int(info) != 0:
        factor = shifted_cholesky(h, len(h) * torch.finfo(h.real.dtype).eps)
        if factor is None:
            return torch.zeros_like(rhs)
    x = torch.cholesky_solve(q.to(factor.dtype), factor).to(q.dtype)
    return x[:, 0] if rhs.ndim == 1 else x


def row_chunk(model, columns, config):
    """Rows per streamed block: at most dictionary_elements / columns, whole apertures."""
    chunk = min(config.row_chunk, max(1, config.dictionary_elements // columns))
    return max(model.row_alignment, chunk - chunk % model.row_alignment)


SYRK_MODE = {"plain": "fp32", "fp32": "fp32", "bf16x3": "bf16x3", "fp16x3": "fp16x3"}


def normal_equations(model, params, target, config):
    """G = A^H A / M and q = A^H y / M (Re of both for real coefficients)."""
    count = len(params["centers"])
    stream = model.blocks(params, target, config.real, row_chunk(model, count, config))
    if not config.real:
        gram = target.new_zeros((count, count))
        rhs = target.new_zeros(count)
        for a, y in stream:
            gram.addmm_(a.H, a)
            rhs.addmv_(a.H, y)
        return (gram + gram.H) / (2 * model.size), rhs / model.size
    mode = SYRK_MODE[getattr(config, "gram", "plain")]
    batched = getattr(config, "batched", False) and hasattr(model, "correlate")
    device = target.device
    part = torch.zeros((count, count), dtype=torch.float32, device=device)
    total = torch.zeros((count, count), dtype=torch.float64, device=device)
    rhs = torch.zeros(count, dtype=torch.float64, device=device)
    blocks = ((a, y, None) for a, y in stream)
    if batched and hasattr(model, "projected_blocks"):
        # projected blocks straight from the atoms on chip (with their column maxima for
        # the FP16 split); y enters through the correlation
        fused = model.projected_blocks(params, row_chunk(model, count, config), column_max=True)
        if fused is not None:
            blocks = ((a, None, maxima) for a, maxima in fused)
    for index, (a, y, maxima) in enumerate(blocks):
        syrk(a.contiguous(), part, mode, maxima)
        if not batched:
            rhs += a.T @ y
        if (index + 1) % FOLD == 0:
            total += part
            part.zero_()
    total += part
    upper = torch.triu(total)
    gram = (upper + torch.triu(total, 1).T) / model.size
    if batched:
        # A^H y in one correlation launch over every aperture and bin: the row basis holds
        # every atom, so projecting y first changes nothing
        rhs = model.correlate(params, target)[0].real.double()
    return gram.float(), (rhs / model.size).float()


def fit_coefficients(gram, rhs, config):
    """Minimizer of b^H G b - 2 Re(b^H q) + ridge |b|^2 in the configured domain."""
    if config.real:
        gram, rhs = gram.real, rhs.real
    eye = torch.eye(len(gram), dtype=gram.dtype, device=gram.device)
    return solve(gram + config.ridge * eye, rhs)


def rendered_loss(model, params, coefficients, target, ridge):
    """Prediction and the complete objective, from a fresh render (varpro.py:132-136)."""
    prediction = model.render(params, coefficients)
    loss = energy(prediction, target) / model.size + ridge * coefficients.abs().square().sum()
    return prediction, float(loss)


@torch.no_grad()
def varpro(model, params, target, config):
    """Coefficients, prediction, objective, Gram and right-hand side of one geometry."""
    gram, rhs = normal_equations(model, params, target, config)
    b = fit_coefficients(gram, rhs, config)
    return Fit(b, *rendered_loss(model, params, b, target, config.ridge), gram, rhs)


@torch.no_grad()
def exchange_varpro(state, proposal, slot, config):
    """VarPro after one atom changed: only its Gram column and rhs entry are recomputed,
    from one rendered column correlated against every atom (varpro.py:147-163)."""
    model, target = state.model, state.target
    atom = {key: value[slot : slot + 1] for key, value in proposal.items()}
    column = model.render(atom, target.real.new_ones(1))
    cross = model.correlate(proposal, column)[0] / model.size
    value = (column.conj() * target).sum() / model.size
    if config.real:
        cross, value = cross.real, value.real
    gram, rhs = state.gram.clone(), state.rhs.clone()
    gram[:, slot], gram[slot, :] = cross, cross.conj()
    gram[slot, slot] = cross[slot].real
    rhs[slot] = value
    b = fit_coefficients(gram, rhs, config)
    return Fit(b, *rendered_loss(model, proposal, b, target, config.ridge), gram, rhs)


def screen_bar(screen, params, config):
    """(model, target, subset objective of params) for an aperture subset, or None."""
    if screen is None:
        return None
    model, target = screen
    return model, target, varpro(model, params, target, config).objective


@torch.no_grad()
def accept(state, proposal, config, screen=None):
    """The state with proposal installed if its refitted objective is lower, else None.

    A screen first rejects proposals clearly worse on its aperture subset; one changed
    atom refits only its Gram column (varpro.py:178-203).
    """
    changed = torch.zeros(len(state.coefficients), dtype=torch.bool, device=state.target.device)
    for key, value in proposal.items():
        changed |= (value != state.params[key]).reshape(len(changed), -1).any(-1)
    slots = changed.nonzero().flatten().tolist()
    if not slots:
        return None
    if screen is not None:
        model, target, bar = screen
        if varpro(model, proposal, target, config).objective > bar + 1e-4 * state.power:
            return None
    if len(slots) == 1:
        fit = exchange_varpro(state, proposal, slots[0], config)
    else:
        fit = varpro(state.model, proposal, state.target, config)
    if not fit.objective < state.objective:
        return None
    return replace(state, params=proposal, **fit._asdict())


def pcg(operator, rhs, diagonal, *, iterations=32, tolerance=1e-3):
    """Jacobi-preconditioned conjugate gradients from zero (varpro.py:206-228)."""
    x = torch.zeros_like(rhs)
    r = rhs.clone()
    target_norm = tolerance * rhs.norm()
    z = r / diagonal
    p = z.clone()
    rz = r @ z
    for _ in range(iterations):
        hp = operator(p)
        curvature = p @ hp
        if not curvature > 0:
            break
        alpha = rz / curvature
        x += alpha * p
        r -= alpha * hp
        if r.norm() <= target_norm:
            break
        z = r / diagonal
        rz_next = r @ z
        p = z + (rz_next / rz) * p
        rz = rz_next
    return x
