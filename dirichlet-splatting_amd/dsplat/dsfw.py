"""DSFW: replace support, refine geometry, accept on the full objective, repeat.

Same interface as dsplat/dsfw.py of the reference (dsfw.py:23-174). Every function
takes a state and returns a new one. Defaults are the fast 3-D configuration:
signed-real coefficients, merged replacement, joint Gauss-Newton on 1024 apertures.
DSFWConfig.gram selects the SYRK of varpro's normal equations ("plain" FP32 MFMA, or the
split-precision "bf16x3" / "fp16x3" MFMA products).
"""

from dataclasses import dataclass, replace

import torch

from .parameters import ParameterConfig
from .refinement import dense_slide, joint_slide, trainable
from .replacement import merge_support, replace_one
from .spatial import propose
from .varpro import varpro

LM_WAIT_LIMIT = 8


@dataclass(frozen=True)
class DSFWConfig:
    iterations: int = 80
    ridge: float = 0.0
    coefficients: str = "real"  # "real" (signed) or "complex"
    replacement: str = "merged"  # "merged" or "utility"
    batch: int = 1024
    min_separation: float = 0.0
    lm_solver: str = "joint_cg"  # "joint_cg" or "dense"
    lm_steps: int = 1
    lm_gain_ratio: float = 1.0
    lm_rows: int = 4096
    joint_apertures: int = 1024
    screen_apertures: int = 1024
    cg_iterations: int = 16
    cg_tolerance: float = 1e-3
    cg_probes: int = 2
    damping: float = 1e-3
    trust_radius: float = 1.5
    row_chunk: int = 65536
    dictionary_elements: int = 32_000_000
    seed: int = 0
    gram: str = "plain"  # normal-equation SYRK: plain (FP32) | bf16x3 | fp16x3
    batched: bool = False  # one-launch candidate scoring, backtracking and elimination

    @property
    def real(self):
        """Signed-real coefficients; a misspelt mode raises KeyError."""
        return {"real": True, "complex": False}[self.coefficients]


@dataclass(frozen=True)
class DSFWState:
    model: object
    target: torch.Tensor
    params: dict
    candidates: dict
    parameters: dict
    coefficients: torch.Tensor
    prediction: torch.Tensor
    objective: float
    gram: torch.Tensor
    rhs: torch.Tensor
    power: float
    iteration: int = 0
    history: tuple = ()
    peak_cache: dict | None = None
    search_failures: int = 0
    lm_skip: int = 0
    lm_wait: int = 0


def _record(state, config, **extra):
    """The state with a history entry; the data term of the objective gives the NMSE."""
    data = state.objective - config.ridge * float(state.coefficients.abs().square().sum())
    nmse = data / max(state.power, torch.finfo(torch.float32).tiny)
    entry = dict(iteration=state.iteration, objective=state.objective, nmse=nmse, **extra)
    return replace(state, history=(*state.history, entry))


@torch.no_grad()
def initialize(model, target, params, config, *, candidates=None, parameters=None):
    """Project params onto their constraints and fit the coefficients (dsfw.py:88-110)."""
    specs = {key: (parameters or {}).get(key, ParameterConfig(trainable=False)) for key in params}
    params = {key: specs[key].project(value.detach().clone(), value.detach()) for key, value in params.items()}
    model = model.prepare(specs)
    candidates = params if candidates is None else {key: v.detach().clone() for key, v in candidates.items()}
    fit = varpro(model, params, target, config)
    state = DSFWState(
        model, target, params, candidates, specs, **fit._asdict(), power=float(target.abs().square().mean())
    )
    return _record(state, config)


@torch.no_grad()
def step(state, config, search=None):
    """One DSFW iteration; state.history[-1] describes it (dsfw.py:113-150)."""
    before = state.objective
    pool, scores = state.candidates, None
    if search is not None:
        pool, scores, cache = propose(state, config, search)
        state = replace(state, peak_cache=cache)
    screen = None
    if 0 < config.screen_apertures < state.model.apertures:
        ids, model = state.model.sample_apertures(config.screen_apertures, config.seed + state.iteration + 2000003)
        screen = model, state.target[ids]
    if {"merged": True, "utility": False}[config.replacement]:
        state, replaced = merge_support(state, config, pool, scores)
    else:
        state, replaced = replace_one(state, config, pool, screen)
    if search is not None:
        state = replace(state, search_failures=0 if replaced else state.search_failures + 1)
    middle = state.objective
    slides, skip, wait = 0, state.lm_skip, state.lm_wait
    if trainable(state.params, state.parameters) and skip == 0:
        slide = {"joint_cg": joint_slide, "dense": dense_slide}[config.lm_solver]
        state, slides = slide(state, config, screen)
        if config.lm_gain_ratio > 0:
            poor = middle - state.objective < config.lm_gain_ratio * (before - middle)
            wait = min(LM_WAIT_LIMIT, max(1, 2 * wait)) if poor else 0
            skip = wait
    elif skip:
        skip -= 1
    state = replace(state, iteration=state.iteration + 1, lm_skip=skip, lm_wait=wait)
    return _record(state, config, replaced=replaced, lm_accepted=slides)


def fit(model, target, params, config=None, search=None, *, candidates=None, parameters=None, callback=None):
    """initialize, then config.iterations steps; callback receives each history entry."""
    config = config or DSFWConfig()
    state = initialize(model, target, params, config, candidates=candidates, parameters=parameters)
    for _ in range(config.iterations):
        state = step(state, config, search)
        if callback is not None:
            callback(state.history[-1])
    return state
