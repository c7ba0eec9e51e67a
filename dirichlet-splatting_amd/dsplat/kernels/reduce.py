"""Binding of the energy reduction (mirrors dsplat/kernels/reduce.py of the reference)."""

import torch

from .. import _lib


def energy(value, reference):
    """sum |value - reference|^2 of two complex tensors in one deterministic reduction
    (kernels/reduce.py:8-18); a 0-d float32 tensor."""
    value = torch.view_as_real(value.to(torch.complex64).resolve_conj().resolve_neg().contiguous()).reshape(-1, 2)
    reference = torch.view_as_real(
        reference.to(torch.complex64).resolve_conj().resolve_neg().contiguous()
    ).reshape(-1, 2)
    if value.shape != reference.shape:
        raise ValueError("energy needs two tensors of one shape")
    out = value.new_empty(())
    size = _lib.library().dsplat_energy_workspace(len(value))
    workspace = value.new_empty(max(1, size // 4))
    _lib.call(
        "dsplat_energy", _lib.ptr(value), _lib.ptr(reference), len(value), _lib.ptr(out), _lib.ptr(workspace), size,
        _lib.stream(value.device),
    )  # fmt: skip
    return out
