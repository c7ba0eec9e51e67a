"""Variable projection: the normal equations of one geometry, the coefficient solve, and the
acceptance test of a proposed geometry.

Public surface of the reference's dsplat/varpro.py:17-228 (Fit, largest_eigenvalue,
shifted_cholesky, solve, row_chunk, normal_equations, fit_coefficients, rendered_loss,
varpro, exchange_varpro, screen_bar, accept, pcg). The objective is
mean(|A b - y|^2) + ridge |b|^2; complex coefficients use G = A^H A / M, signed real ones
Re(G) and Re(q). A geometry is accepted only when its refitted, freshly rendered objective is
lower.

The real normal equations run on the MFMA products of kernels.blas, chosen by DSFWConfig.gram:

  plain   float32 row-projected blocks (design, then the range subspace), FP32 SYRK: the
          reference's stream over cuBLAS SSYRK
  bf16x3  with the batched search on a sample-domain chirp: blocks written by the projection
          kernel itself as BF16 hi / lo operands, one split SYRK per block (hi hi + hi lo +
          lo hi); otherwise the float32 blocks split inside the SYRK
  fp16x3  the same in FP16 (22 bits per operand instead of 17), every column scaled by a
          power of two from an a-priori bound of its entries so that FP16 holds it;
          otherwise float32 blocks with their column maxima, split inside the SYRK

Block products accumulate in float32 and fold into a float64 total every FOLD blocks.
"""

from dataclasses import replace
from typing import NamedTuple

import torch

from .kernels.blas import gemm_tn, syrk, syrk_split
from .kernels.reduce import energy
from .kernels.solver import cholesky_ex, cholesky_solve, power_iteration


class Fit(NamedTuple):
    """The VarPro solution of one geometry; the field names are DSFWState's."""

    coefficients: torch.Tensor
    prediction: torch.Tensor
    objective: float
    gram: torch.Tensor  # Re(A^H A) / M for real coefficients, A^H A / M otherwise
    rhs: torch.Tensor


POWER_STEPS = 24  # power iterations of the largest-eigenvalue estimate
SHIFT_RETRIES = 4  # diagonal shifts tried before a system counts as empty
FOLD = 8  # float32 block products summed before each float64 fold

SYRK_MODE = {"plain": "fp32", "fp32": "fp32", "bf16x3": "bf16x3", "fp16x3": "fp16x3"}


# --- Solves --------------------------------------------------------------------------------


def largest_eigenvalue(h):
    """Power-iteration estimate of lambda_max of a positive semidefinite h, from a fixed
    non-constant start (fractional parts of multiples of the golden ratio, plus 1/2), so that
    the estimate is reproducible."""
    index = torch.arange(len(h), device=h.device, dtype=torch.float64)
    vector = (torch.remainder(index * 0.6180339887, 1) + 0.5).to(h.dtype)
    last = power_iteration(h, vector, POWER_STEPS)  # real float64 on the device: one launch per step
    if last is not None:
        return last.norm()
    top = vector.norm()
    for _ in range(POWER_STEPS):
        vector = h @ vector
        top = vector.norm()
        vector = vector / top.clamp(min=torch.finfo(top.dtype).tiny)
    return top


def shifted_cholesky(h, floor):
    """Cholesky factor of h + shift I in double precision, or None. The shift starts at floor
    lambda_max(h) and is added again, nine times larger each time, while the factorization
    fails (1, 10, 91, ... times the first), at most SHIFT_RETRIES times."""
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
    """Hermitian solve h x = rhs. A singular system (a column no aperture sees, two atoms in
    one place) fails the first Cholesky and is shifted to the level where a pseudoinverse
    truncates, len(h) eps lambda_max: that moves x only along near-null directions, which do
    not change the prediction. An all-zero system gives zeros."""
    columns = rhs[:, None] if rhs.ndim == 1 else rhs
    factor, info = cholesky_ex(h)
    if int(info) != 0:
        factor = shifted_cholesky(h, len(h) * torch.finfo(h.real.dtype).eps)
        if factor is None:
            return torch.zeros_like(rhs)
    x = cholesky_solve(columns.to(factor.dtype), factor).to(columns.dtype)
    return x[:, 0] if rhs.ndim == 1 else x


def fit_coefficients(gram, rhs, config):
    """The minimizer of b^H G b - 2 Re(b^H q) + ridge |b|^2 (real b for real coefficients)."""
    if config.real:
        gram, rhs = gram.real, rhs.real
    return solve(gram + config.ridge * torch.eye(len(gram), dtype=gram.dtype, device=gram.device), rhs)


# --- Normal equations ----------------------------------------------------------------------


def row_chunk(model, columns, config):
    """Observation rows per streamed block: at most dictionary_elements / columns and
    row_chunk, whole apertures."""
    rows = min(config.row_chunk, max(1, config.dictionary_elements // columns))
    return max(model.row_alignment, rows - rows % model.row_alignment)


class _Folded:
    """A float32 accumulator [rows, cols] that folds into a float64 total every FOLD blocks."""

    def __init__(self, rows, cols, device):
        self.part = torch.zeros((rows, cols), dtype=torch.float32, device=device)
        self.total = torch.zeros((rows, cols), dtype=torch.float64, device=device)
        self.blocks = 0

    def added(self):
        """Call after each block's product went into self.part."""
        self.blocks += 1
        if self.blocks % FOLD == 0:
            self.total += self.part
            self.part.zero_()

    def result(self):
        self.total += self.part
        self.part.zero_()
        return self.total


def block_products(model, params, target, config, out, col0=None, known=None):
    """Accumulate the real block products of the streamed design into `out` (a _Folded of
    [S, S] upper triangle, or [S, S - col0] cross products of the trailing columns) with the
    configured backend. Returns the float64 A^T y of the streamed blocks, or None for the cross
    products and in the batched form (which takes A^H y from one correlation). known: the
    leading atoms of params whose projections a model's split cache may already hold."""
    count = len(params["centers"])
    chunk = row_chunk(model, count, config)
    mode = SYRK_MODE[getattr(config, "gram", "plain")]
    batched = getattr(config, "batched", False) and hasattr(model, "adjoint_correlation")
    if batched and mode in ("bf16x3", "fp16x3") and hasattr(model, "split_blocks"):
        split = model.split_blocks(params, chunk, {"bf16x3": "bf16", "fp16x3": "fp16"}[mode], known=known)
        if split is not None:
            for block in split:
                syrk_split(block, out.part, col0)
                out.added()
            return None
    if batched and hasattr(model, "projected_blocks"):
        fused = model.projected_blocks(params, chunk, column_max=True)
        if fused is not None:
            for a, maxima in fused:
                _block_product(a, out.part, mode, maxima, col0)
                out.added()
            return None
    rhs = None if batched or col0 is not None else torch.zeros(count, dtype=torch.float64, device=target.device)
    for a, y in model.blocks(params, target, True, chunk):
        _block_product(a.contiguous(), out.part, mode, None, col0)
        if rhs is not None:
            rhs += a.T @ y
        out.added()
    return rhs


def _block_product(a, out, mode, maxima, col0):
    if col0 is None:
        syrk(a.contiguous(), out, mode, maxima)
    else:
        gemm_tn(a.contiguous(), col0, out, mode, maxima)


def normal_equations(model, params, target, config):
    """G = A^H A / M and q = A^H y / M (their real parts for real coefficients), streamed over
    row-projected blocks of whole apertures (varpro.py:91-120 of the reference)."""
    count = len(params["centers"])
    if not config.real:
        gram = target.new_zeros((count, count))
        rhs = target.new_zeros(count)
        for a, y in model.blocks(params, target, False, row_chunk(model, count, config)):
            gram.addmm_(a.H, a)
            rhs.addmv_(a.H, y)
        return (gram + gram.H) / (2 * model.size), rhs / model.size
    acc = _Folded(count, count, target.device)
    rhs = block_products(model, params, target, config, acc)
    total = acc.result()
    gram = (torch.triu(total) + torch.triu(total, 1).T) / model.size
    if rhs is None:
        # A^H y by one adjoint over every aperture and bin: the range subspace holds every
        # atom, so projecting y first changes nothing
        rhs = model.adjoint_correlation(params, target).real.double()
    return gram.float(), (rhs / model.size).float()


# --- Objective and acceptance --------------------------------------------------------------


def rendered_loss(model, params, coefficients, target, ridge):
    """The prediction and the complete objective, from a fresh render (varpro.py:132-136)."""
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
    """VarPro after one atom moved: its Gram row and column and its rhs entry come from one
    rendered column correlated against every atom; the rest is the state's
    (varpro.py:147-163)."""
    model, target = state.model, state.target
    column = model.render({key: value[slot : slot + 1] for key, value in proposal.items()}, target.real.new_ones(1))
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
    """(model, target, subset objective of params) of an aperture subset, or None: one bar
    per state, against which every proposal from that state is screened."""
    if screen is None:
        return None
    model, target = screen
    return model, target, varpro(model, params, target, config).objective


def _changed_atoms(state, proposal):
    moved = torch.zeros(len(state.coefficients), dtype=torch.bool, device=state.target.device)
    for key, value in proposal.items():
        moved |= (value != state.params[key]).reshape(len(moved), -1).any(-1)
    return moved.nonzero().flatten().tolist()


@torch.no_grad()
def accept(state, proposal, config, screen=None):
    """The state with proposal installed if its refitted objective is lower, else None. A
    screen (screen_bar of state.params) first turns away proposals clearly worse on its
    aperture subset; a single moved atom refits only its Gram column (varpro.py:178-203)."""
    slots = _changed_atoms(state, proposal)
    if not slots:
        return None
    if screen is not None:
        model, target, bar = screen
        if varpro(model, proposal, target, config).objective > bar + 1e-4 * state.power:
            return None
    fit = (exchange_varpro(state, proposal, slots[0], config) if len(slots) == 1
           else varpro(state.model, proposal, state.target, config))  # fmt: skip
    if not fit.objective < state.objective:
        return None
    return replace(state, params=proposal, **fit._asdict())


def pcg(operator, rhs, diagonal, *, iterations=32, tolerance=1e-3, fused=False, graph=False):
    """Jacobi-preconditioned conjugate gradients from zero: stops at the first direction of
    non-positive curvature or once the residual is tolerance times the right-hand side
    (varpro.py:206-228).

    The two stopping tests stay on the device: a flag turns every later update into a no-op
    (torch.where, so values after a stop, NaN included, never leak), and the loop always runs
    `iterations` operator applications. The result equals the early-exit loop's bit for bit,
    and nothing waits for the host between iterations.

    fused runs each iteration's vector algebra as one kernel (float32 vectors, float64 sums,
    on float32 device vectors); graph then also replays the operator and that kernel from a
    HIP graph captured once per call (the operator must launch only on the current stream,
    allocate only through torch and never synchronize)."""
    if fused:
        return _pcg_device(operator, rhs, diagonal, iterations, tolerance, graph)
    x = torch.zeros_like(rhs)
    residual = rhs.clone()
    goal = tolerance * rhs.norm()
    z = residual / diagonal
    direction = z.clone()
    rz = residual @ z
    active = torch.ones((), dtype=torch.bool, device=rhs.device)
    for _ in range(iterations):
        image = operator(direction)
        curvature = direction @ image
        moving = active & (curvature > 0)
        step = torch.where(moving, rz / curvature, 0)
        x = torch.where(moving, x + step * direction, x)
        residual = torch.where(moving, residual - step * image, residual)
        active = moving & ~(residual.norm() <= goal)
        z = residual / diagonal
        rz_next = residual @ z
        direction = torch.where(active, z + (rz_next / rz) * direction, direction)
        rz = torch.where(active, rz_next, rz)
    return x


def _pcg_device(operator, rhs, diagonal, iterations, tolerance, graph):
    """pcg on the fused iteration kernel: the first iteration eagerly, then, with graph, the
    next one captured on a side stream and replayed for the rest."""
    from .kernels.solver import pcg_start, pcg_step

    diagonal = diagonal.float().contiguous()
    x, residual, direction, state = pcg_start(rhs, diagonal, tolerance)

    def iteration():
        pcg_step(operator(direction), diagonal, x, residual, direction, state)

    if iterations < 1:
        return x
    iteration()
    if not graph or iterations == 1:
        for _ in range(iterations - 1):
            iteration()
        return x
    device = rhs.device
    side = torch.cuda.Stream(device)
    side.wait_stream(torch.cuda.current_stream(device))
    captured = torch.cuda.CUDAGraph()
    with torch.cuda.stream(side):
        captured.capture_begin()
        try:
            iteration()
        finally:
            captured.capture_end()
    torch.cuda.current_stream(device).wait_stream(side)
    for _ in range(iterations - 1):
        captured.replay()
    return x
