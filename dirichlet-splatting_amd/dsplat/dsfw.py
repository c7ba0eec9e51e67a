"""DSFW, the solver loop: replace part of the support, refine the geometry, accept only what
lowers the full objective, repeat.

Public surface of the reference's dsplat/dsfw.py:23-174 (DSFWConfig, DSFWState, initialize,
step, fit). States are immutable: every function returns a new one, sharing tensors. The
defaults are the fast 3-D configuration: signed real coefficients, merged replacement of up
to `batch` candidates per step, and joint Gauss-Newton on an aperture subset that pauses
while it gains less than the replacement.

Two fields select this build's MI355X paths (their defaults are the reference's arithmetic):
`gram` the MFMA backend of the real normal equations (see varpro), and `batched` the
one-launch orientation choice, quasi-Newton search and elimination, with right-hand sides
from one correlation launch.
"""

from dataclasses import dataclass, replace

import torch

from .parameters import ParameterConfig
from .refinement import dense_slide, joint_slide, trainable
from .replacement import merge_support, replace_one
from .spatial import propose
from .varpro import varpro

# Longest pause of the geometry step, in steps: a pause starts at one step and doubles while
# the geometry step keeps gaining less than lm_gain_ratio of the replacement's gain.
LM_WAIT_LIMIT = 8


@dataclass(frozen=True)
class DSFWConfig:
    iterations: int = 80
    ridge: float = 0.0  # on the mean-square data term
    coefficients: str = "real"  # "real" (signed) or "complex"
    replacement: str = "merged"  # "merged" or "utility" (one slot per step)
    batch: int = 1024  # merged: candidates pooled with the support per step
    min_separation: float = 0.0  # replacement candidates' distance from the support
    lm_solver: str = "joint_cg"  # "joint_cg" or "dense"
    lm_steps: int = 1
    lm_gain_ratio: float = 1.0  # the geometry step pauses while it gains less than this share
    lm_rows: int = 4096  # dense: sampled observation rows
    joint_apertures: int = 1024  # apertures of the joint_cg system (0 or >= all: every one)
    screen_apertures: int = 1024  # apertures of the quick rejection of worse proposals
    cg_iterations: int = 16
    cg_tolerance: float = 1e-3
    cg_probes: int = 2
    damping: float = 1e-3
    trust_radius: float = 1.5  # dimensionless: times each ParameterConfig.scale
    row_chunk: int = 65536
    dictionary_elements: int = 32_000_000  # observations x atoms per streamed block
    seed: int = 0
    gram: str = "plain"  # real normal equations: plain (FP32) | bf16x3 | fp16x3 split products
    batched: bool = False  # one-launch scoring, search and elimination

    @property
    def real(self):
        """Signed real coefficients; a misspelt mode raises KeyError."""
        return {"real": True, "complex": False}[self.coefficients]


@dataclass(frozen=True)
class DSFWState:
    model: object
    target: torch.Tensor
    params: dict
    candidates: dict  # the replacement pool when there is no spatial search
    parameters: dict  # name -> ParameterConfig
    coefficients: torch.Tensor
    prediction: torch.Tensor
    objective: float
    gram: torch.Tensor  # Re(A^H A) / M for real coefficients, A^H A / M otherwise
    rhs: torch.Tensor
    power: float  # mean |target|^2
    iteration: int = 0
    history: tuple = ()
    peak_cache: dict | None = None  # spatial search peaks carried to the next step
    search_failures: int = 0
    lm_skip: int = 0  # steps left before the geometry step runs again
    lm_wait: int = 0  # length of the current pause


def _logged(state, config, **extra):
    """state with a history entry for its current objective; the data term over the target
    power is the NMSE (0 for an exactly fitted zero target)."""
    data_term = state.objective - config.ridge * float(state.coefficients.abs().square().sum())
    nmse = data_term / max(state.power, torch.finfo(torch.float32).tiny)
    entry = dict(iteration=state.iteration, objective=state.objective, nmse=nmse, **extra)
    return replace(state, history=(*state.history, entry))


@torch.no_grad()
def initialize(model, target, params, config, *, candidates=None, parameters=None):
    """The first state: params projected onto their constraints (fields without a
    ParameterConfig stay fixed) and their coefficients fitted. candidates (default: params)
    form the replacement pool without a spatial search; fields they omit copy slot 0 (merged)
    or the replaced slot (utility)."""
    specs = {key: (parameters or {}).get(key, ParameterConfig(trainable=False)) for key in params}
    params = {key: specs[key].project(value.detach().clone(), value.detach()) for key, value in params.items()}
    model = model.prepare(specs)
    pool = params if candidates is None else {key: value.detach().clone() for key, value in candidates.items()}
    solution = varpro(model, params, target, config)
    state = DSFWState(model, target, params, pool, specs, **solution._asdict(),
                      power=float(target.abs().square().mean()))  # fmt: skip
    return _logged(state, config)


def _screen(state, config):
    """The seeded aperture subset (model, target) that screens proposals, or None."""
    if not 0 < config.screen_apertures < state.model.apertures:
        return None
    ids, model = state.model.sample_apertures(config.screen_apertures, config.seed + state.iteration + 2000003)
    return model, state.target[ids]


def _geometry_step(state, config, screen, before, middle):
    """The refinement step unless it is paused: (state, accepted steps, pause left, pause
    length). A step that gains less than lm_gain_ratio of the replacement's gain starts or
    doubles a pause; one that pays ends it."""
    skip, wait = state.lm_skip, state.lm_wait
    if skip or not trainable(state.params, state.parameters):
        return state, 0, max(0, skip - 1), wait
    slide = {"joint_cg": joint_slide, "dense": dense_slide}[config.lm_solver]
    state, accepted = slide(state, config, screen)
    if config.lm_gain_ratio > 0:
        poor = middle - state.objective < config.lm_gain_ratio * (before - middle)
        wait = min(LM_WAIT_LIMIT, max(1, 2 * wait)) if poor else 0
        skip = wait
    return state, accepted, skip, wait


@torch.no_grad()
def step(state, config, search=None):
    """One DSFW iteration; state.history[-1] describes it. With a SpatialSearchConfig the
    replacement pool is a fresh spatial search (dsfw.py:113-150)."""
    before = state.objective
    pool, scores = state.candidates, None
    if search is not None:
        pool, scores, cache = propose(state, config, search)
        state = replace(state, peak_cache=cache)
    screen = _screen(state, config)
    if {"merged": True, "utility": False}[config.replacement]:
        state, replaced = merge_support(state, config, pool, scores)
    else:
        state, replaced = replace_one(state, config, pool, screen)
    if search is not None:
        state = replace(state, search_failures=0 if replaced else state.search_failures + 1)
    state, accepted, skip, wait = _geometry_step(state, config, screen, before, state.objective)
    state = replace(state, iteration=state.iteration + 1, lm_skip=skip, lm_wait=wait)
    return _logged(state, config, replaced=replaced, lm_accepted=accepted)


def fit(model, target, params, config=None, search=None, *, candidates=None, parameters=None, callback=None):
    """initialize, then config.iterations steps; callback receives each step's history entry."""
    config = config or DSFWConfig()
    state = initialize(model, target, params, config, candidates=candidates, parameters=parameters)
    for _ in range(config.iterations):
        state = step(state, config, search)
        if callback is not None:
            callback(state.history[-1])
    return state
