"""Bindings of the search kernels: range gating, separated peaks and quasi-Newton steps
(mirrors dsplat/kernels/search.py of the reference)."""

import math

import torch

from .. import _lib
from . import interleaved


def range_gate(points, positions, intervals, spacing, min_fraction):
    """Points whose spherical range bin lies in a marked interval at min_fraction of apertures
    (kernels/search.py:10-24)."""
    output = torch.empty(len(points), device=points.device, dtype=torch.int32)
    required = math.ceil(min_fraction * len(positions) - 1e-7)
    _lib.call(
        "dsplat_search_range_gate", _lib.ptr(points.float().contiguous()), len(points),
        _lib.ptr(positions.float().contiguous()), len(positions), _lib.ptr(intervals.to(torch.int32).contiguous()),
        intervals.shape[1], float(spacing), required, _lib.ptr(output), _lib.stream(points.device),
    )  # fmt: skip
    return output.bool()


def peaks(scores, centers, count, separation):
    """Greedy best-first indices at least separation apart; scores are not modified
    (kernels/search.py:27-39). The kernel sorts once and resolves blocks of sorted
    candidates instead of one arg-max sweep per peak."""
    output = torch.full((min(count, len(scores)),), -1, device=centers.device, dtype=torch.int32)
    if len(output) == 0:
        return output.long()
    centers = centers.float().contiguous()
    dim = centers.shape[1]
    size = _lib.library().dsplat_search_peaks_workspace(len(scores), len(output), dim)
    workspace = torch.empty(max(1, size), dtype=torch.uint8, device=centers.device)
    _lib.call(
        "dsplat_search_peaks", _lib.ptr(scores.float().contiguous()), _lib.ptr(centers), len(scores), dim,
        len(output), float(separation), _lib.ptr(output), _lib.ptr(workspace), size, _lib.stream(centers.device),
    )  # fmt: skip
    return output[output >= 0].long()


TOP_MAX = 4096  # DSPLAT_TOP_MAX


def certificate_scores(rows, areas, power, real_part):
    """float64 certificates [n, C] of scan rows [n, C, 3] (corr.re, corr.im, norm) with area:
    |s|^2 area / max(norm area, 1e-60) / power (certificate.py:11-31)."""
    n, count = rows.shape[:2]
    scores = torch.empty((n, count), dtype=torch.float64, device=rows.device)
    _lib.call(
        "dsplat_certificate_scores", _lib.ptr(rows.float().contiguous()), n, count,
        _lib.ptr(areas.float().contiguous()), _lib.ptr(power.double().contiguous()), int(real_part),
        _lib.ptr(scores), _lib.stream(rows.device),
    )  # fmt: skip
    return scores


def top_descending(scores, top):
    """Indices [rows, min(top, count)] of each row's largest float64 scores [rows, count],
    descending, ties to the lower index: argsort(descending=True, stable=True)[:, :top] by one
    selection launch."""
    rows, count = scores.shape
    if not 1 <= top <= TOP_MAX:
        raise ValueError(f"top must lie in [1, {TOP_MAX}], got {top}")
    out = torch.empty((rows, min(top, count)), dtype=torch.int64, device=scores.device)
    _lib.call(
        "dsplat_top_descending", _lib.ptr(scores.double().contiguous()), rows, count, int(top), _lib.ptr(out),
        _lib.stream(scores.device),
    )  # fmt: skip
    return out


def direction(inverse, gradient, normals, mask, radius):
    """Quasi-Newton descent per candidate: tangent to unit normals, inside the trust radius
    (kernels/search.py:42-56)."""
    count, width = gradient.shape
    output = torch.empty_like(gradient)
    _lib.call(
        "dsplat_search_direction", _lib.ptr(inverse.contiguous()), _lib.ptr(gradient.contiguous()),
        _lib.ptr(normals.contiguous()), _lib.ptr(mask.float().contiguous()), count, width, float(radius),
        _lib.ptr(output), _lib.stream(gradient.device),
    )  # fmt: skip
    return output


def trial(centers, normals, direction, controls, fraction):
    """Step fraction along direction, clamped to the box and renormalized; frozen axes kept.

    controls rows: per-coordinate scale, trainable mask, lower and upper bounds
    (kernels/search.py:59-77).
    """
    count, width = direction.shape
    out_centers, out_normals = torch.empty_like(centers), torch.empty_like(normals)
    _lib.call(
        "dsplat_search_trial", _lib.ptr(centers.contiguous()), _lib.ptr(normals.contiguous()),
        _lib.ptr(direction.contiguous()), _lib.ptr(controls.float().contiguous()), count, width, float(fraction),
        _lib.ptr(out_centers), _lib.ptr(out_normals), _lib.stream(centers.device),
    )  # fmt: skip
    return out_centers, out_normals


def qn_search(config, positions, residual, centers, normals, areas, spec, fresnel_n):
    """Batched quasi-Newton certificate ascent of every candidate in one launch: final
    (centers, normals, values) (certificate.py:106-201). spec is a filled _lib.QnSpec."""
    count = len(centers)
    out_c, out_n = torch.empty_like(centers), torch.empty_like(normals)
    values = torch.empty(count, dtype=torch.float32, device=centers.device)
    if count == 0:
        return out_c, out_n, values
    cfg = _lib.fmcw(config, fresnel_n)
    size = _lib.library().dsplat_qn_workspace(_lib.ctypes.byref(cfg), len(positions))
    workspace = torch.empty(max(1, size // 4), dtype=torch.float32, device=centers.device)
    res = interleaved(residual.reshape(len(positions), config.num_range_bins))
    _lib.call(
        "dsplat_qn_search", _lib.ctypes.byref(cfg), _lib.ptr(positions.float().contiguous()), len(positions),
        _lib.ptr(res), _lib.ptr(centers.float().contiguous()), _lib.ptr(normals.float().contiguous()),
        _lib.ptr(areas.float().contiguous()), count, _lib.ctypes.byref(spec), _lib.ptr(out_c), _lib.ptr(out_n),
        _lib.ptr(values), _lib.ptr(workspace), size, _lib.stream(centers.device),
    )  # fmt: skip
    return out_c, out_n, values


def qn_lockstep(positions, centers, normals, areas, spec, operands):
    """The same ascent as qn_search in lockstep rounds of matrix-core evaluations against the
    sample-domain operands of the residual (surfel.CertificateOperands): final (centers,
    normals, values) (certificate.py:106-201)."""
    count = len(centers)
    out_c, out_n = torch.empty_like(centers), torch.empty_like(normals)
    values = torch.empty(count, dtype=torch.float32, device=centers.device)
    if count == 0:
        return out_c, out_n, values
    lib, cfg = _lib.library(), operands.cfg
    size = lib.dsplat_qn_lockstep_workspace(_lib.ctypes.byref(cfg), len(positions), count)
    workspace = torch.empty(max(1, size // 4), dtype=torch.float32, device=centers.device)
    _lib.call(
        "dsplat_qn_lockstep", _lib.ctypes.byref(cfg), _lib.ptr(positions.float().contiguous()), len(positions),
        _lib.ptr(operands.buffer), _lib.ptr(centers.float().contiguous()), _lib.ptr(normals.float().contiguous()),
        _lib.ptr(areas.float().contiguous()), count, _lib.ctypes.byref(spec), _lib.ptr(out_c), _lib.ptr(out_n),
        _lib.ptr(values), _lib.ptr(workspace), size, _lib.stream(centers.device),
    )  # fmt: skip
    return out_c, out_n, values
