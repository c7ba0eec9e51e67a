"""A small launch-shape autotuner: time each candidate shape once with HIP events on the
launch stream and keep the fastest, per (kernel, shape key, device arch).

Results live in memory and, when DSPLAT_AUTOTUNE_CACHE names a writable file, in that
JSON file, so later processes reuse them.
"""

import json
import os
from pathlib import Path

import torch

_MEMORY: dict = {}


def _cache_path():
    path = os.environ.get("DSPLAT_AUTOTUNE_CACHE")
    return Path(path) if path else None


def _load():
    path = _cache_path()
    if path and path.exists():
        try:
            _MEMORY.update(json.loads(path.read_text()))
        except (OSError, ValueError):
            pass


def _store():
    path = _cache_path()
    if path:
        try:
            path.parent.mkdir(parents=True, exist_ok=True)
            path.write_text(json.dumps(_MEMORY, indent=1, sort_keys=True))
        except OSError:
            pass


def tune(kernel, key, candidates, run, device, repeats=2):
    """The fastest of `candidates` for run(candidate); run launches the kernel once."""
    arch = torch.cuda.get_device_properties(device).gcnArchName if torch.cuda.is_available() else "none"
    name = f"{kernel}|{key}|{arch}"
    if not _MEMORY:
        _load()
    if name in _MEMORY:
        return _MEMORY[name]["best"]
    stream = torch.cuda.current_stream(device)
    times = {}
    for candidate in candidates:
        run(candidate)  # warm-up (and code-object load)
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record(stream)
        for _ in range(repeats):
            run(candidate)
        stop.record(stream)
        stop.synchronize()
        times[str(candidate)] = start.elapsed_time(stop) / repeats
    best = min(candidates, key=lambda cand: times[str(cand)])
    _MEMORY[name] = {"best": best, "ms": times}
    _store()
    return best


def record(kernel):
    """The tuning records of a kernel: {shape key: {best, ms per candidate}}."""
    return {k[len(kernel) + 1 : k.rindex("|")]: v for k, v in _MEMORY.items() if k.startswith(kernel + "|")}
