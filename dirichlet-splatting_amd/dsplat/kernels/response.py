"""Binding of the float64 range responses (mirrors dsplat/kernels/response.py of the reference)."""

import torch

from .. import _lib


def range_response(config, low, high, count, device):
    """Float64 range responses [bins, count] of count peak positions spread over [low, high]
    (kernels/response.py:8-20)."""
    output = torch.empty((config.num_range_bins, count, 2), device=device, dtype=torch.float64)
    _lib.call(
        "dsplat_range_response", _lib.ptr(output), float(low), float(high), int(count), config.num_range_bins,
        config.adc_samples, config.n_fft, _lib.stream(torch.device(device)),
    )  # fmt: skip
    return torch.view_as_complex(output)
