"""Bindings of the surfel kernels: oriented surfels seen as FMCW range profiles.

Mirrors dsplat/kernels/surfel.py of the reference (same functions, arguments and
return shapes). config is an FMCWConfig, positions the sensor positions [A, 3].
weights are per-surfel complex amplitudes, reflectivity times sqrt(area); the area
chain lives in dsplat.surfel.
"""

import torch

from .. import _lib
from . import interleaved


def _check(positions, centers, normals):
    if positions.ndim != 2 or positions.shape[1] != 3:
        raise ValueError(f"positions must be [A, 3], got {tuple(positions.shape)}")
    if centers.ndim != 2 or centers.shape[1] != 3 or normals.shape != centers.shape:
        raise ValueError("centers and normals must both be [S, 3]")
    for t in (positions, centers, normals):
        if t.dtype != torch.float32:
            raise TypeError("surfel kernels take float32 geometry")


def _profiles(config, positions, centers, normals, weights, fresnel_n):
    """Range profiles [A, K, 2] = sum_s weights_s * column_s (kernels/surfel.py:58-108)."""
    _check(positions, centers, normals)
    positions, centers, normals = positions.contiguous(), centers.contiguous(), normals.contiguous()
    weights = weights.contiguous()
    output = centers.new_empty((len(positions), config.num_range_bins, 2))
    cfg = _lib.fmcw(config, fresnel_n)
    _lib.call(
        "dsplat_surfel_render",
        _lib.ctypes.byref(cfg),
        _lib.ptr(positions),
        len(positions),
        _lib.ptr(centers),
        _lib.ptr(normals),
        _lib.ptr(weights),
        len(centers),
        _lib.ptr(output),
        _lib.stream(centers.device),
    )
    return output


def vjp(config, positions, centers, normals, weights, upstream, fresnel_n):
    """Gradients of Re(conj(upstream) . profiles) for centers, normals and complex weights
    (kernels/surfel.py:123-137)."""
    _check(positions, centers, normals)
    positions, centers, normals = positions.contiguous(), centers.contiguous(), normals.contiguous()
    count = len(centers)
    w = interleaved(weights)
    up = interleaved(upstream.reshape(len(positions), config.num_range_bins))
    gc = centers.new_empty((count, 3))
    gn = centers.new_empty((count, 3))
    gw = centers.new_empty((count, 2))
    lib = _lib.library()
    cfg = _lib.fmcw(config, fresnel_n)
    size = lib.dsplat_surfel_vjp_workspace(_lib.ctypes.byref(cfg), len(positions), count)
    workspace = centers.new_empty(max(1, size // 4))
    _lib.call(
        "dsplat_surfel_vjp",
        _lib.ctypes.byref(cfg),
        _lib.ptr(positions),
        len(positions),
        _lib.ptr(centers),
        _lib.ptr(normals),
        _lib.ptr(w),
        count,
        _lib.ptr(up),
        _lib.ptr(gc),
        _lib.ptr(gn),
        _lib.ptr(gw),
        _lib.ptr(workspace),
        size,
        _lib.stream(centers.device),
    )
    return gc, gn, torch.view_as_complex(gw)


def jvp(config, positions, centers, normals, weights, delta_centers, delta_normals, delta_weights, fresnel_n):
    """Directional derivative of the profiles [A, K] (kernels/surfel.py:140-164)."""
    _check(positions, centers, normals)
    positions, centers, normals = positions.contiguous(), centers.contiguous(), normals.contiguous()
    out = centers.new_empty((len(positions), config.num_range_bins, 2))
    cfg = _lib.fmcw(config, fresnel_n)
    _lib.call(
        "dsplat_surfel_jvp",
        _lib.ctypes.byref(cfg),
        _lib.ptr(positions),
        len(positions),
        _lib.ptr(centers),
        _lib.ptr(normals),
        _lib.ptr(interleaved(weights)),
        _lib.ptr(delta_centers.to(torch.float32).contiguous()),
        _lib.ptr(delta_normals.to(torch.float32).contiguous()),
        _lib.ptr(interleaved(delta_weights)),
        len(centers),
        _lib.ptr(out),
        _lib.stream(centers.device),
    )
    return torch.view_as_complex(out)


class _Render(torch.autograd.Function):
    """Forward = render; backward = the VJP recomputed from the saved inputs (the forward
    is never stored), as kernels/surfel.py:167-183. jvp gives forward-mode AD through
    the JVP kernel (the reference defines no forward mode)."""

    @staticmethod
    def forward(ctx, centers, normals, weights, positions, config, fresnel_n):
        ctx.save_for_backward(centers, normals, weights, positions)
        ctx.save_for_forward(centers, normals, weights, positions)
        ctx.config, ctx.fresnel_n = config, fresnel_n
        return _profiles(config, positions, centers, normals, weights, fresnel_n)

    @staticmethod
    def jvp(ctx, dcenters, dnormals, dweights, _dpositions, _config, _fresnel):
        # sensor positions carry no tangent: the scan geometry is data, not a parameter
        centers, normals, weights, positions = ctx.saved_tensors
        zero = torch.zeros_like(centers)
        dcenters = zero if dcenters is None else dcenters
        dnormals = zero if dnormals is None else dnormals
        dw = torch.zeros_like(weights) if dweights is None else dweights
        out = jvp(
            ctx.config, positions, centers, normals, torch.view_as_complex(weights), dcenters, dnormals,
            torch.view_as_complex(dw.contiguous()), ctx.fresnel_n,
        )  # fmt: skip
        return torch.view_as_real(out)

    @staticmethod
    def backward(ctx, grad):
        centers, normals, weights, positions = ctx.saved_tensors
        upstream = torch.view_as_complex(grad.contiguous())
        gc, gn, gw = vjp(
            ctx.config, positions, centers, normals, torch.view_as_complex(weights), upstream, ctx.fresnel_n
        )
        return gc, gn, torch.view_as_real(gw), None, None, None


def render(config, positions, centers, normals, weights, fresnel_n):
    """Range profiles [A, K]; differentiable in centers, normals and weights."""
    output = _Render.apply(
        centers.contiguous(), normals.contiguous(), interleaved(weights), positions, config, fresnel_n
    )
    return torch.view_as_complex(output)


def synthesize(config, positions, centers, normals, amplitudes, fresnel_n):
    """Float64 scan [A, K] complex128 of surfels with complex amplitudes by the ADC signal
    chain (groundtruth.py:58-90): beat tones, zero-padded FFT, positive bins, / adc_samples."""
    _check(positions, centers, normals)
    output = torch.empty((len(positions), config.num_range_bins), dtype=torch.complex128, device=centers.device)
    cfg = _lib.fmcw(config, fresnel_n)
    _lib.call(
        "dsplat_synthesize",
        _lib.ctypes.byref(cfg),
        _lib.ptr(positions.contiguous()),
        len(positions),
        _lib.ptr(centers.contiguous()),
        _lib.ptr(normals.contiguous()),
        _lib.ptr(torch.view_as_real(amplitudes.to(torch.complex128).contiguous())),
        len(centers),
        _lib.ptr(torch.view_as_real(output)),
        _lib.stream(centers.device),
    )
    return output


def design(config, positions, centers, normals, areas, fresnel_n, real):
    """Dictionary rows of whole apertures, including sqrt(area) (kernels/surfel.py:199-220).

    Complex rows are [A * K, S]. real gives [2 * A * K, S] ordered bin by bin, row
    (2 k + part) * A + a, so one product projects every aperture onto a row basis.
    """
    _check(positions, centers, normals)
    count, rows = len(centers), len(positions) * config.num_range_bins
    output = centers.new_empty((2 * rows, count) if real else (rows, 2 * count))
    cfg = _lib.fmcw(config, fresnel_n)
    _lib.call(
        "dsplat_surfel_design",
        _lib.ctypes.byref(cfg),
        _lib.ptr(positions.contiguous()),
        len(positions),
        _lib.ptr(centers.contiguous()),
        _lib.ptr(normals.contiguous()),
        _lib.ptr(areas.to(torch.float32).contiguous()),
        count,
        1 if real else 0,
        _lib.ptr(output),
        _lib.stream(centers.device),
    )
    return output if real else torch.view_as_complex(output.view(rows, count, 2))


def jacobian(config, positions, centers, normals, fresnel_n):
    """Columns [A, K, S] and their center and normal derivatives [A, K, S, 3], without area
    (kernels/surfel.py:223-239)."""
    _check(positions, centers, normals)
    out = centers.new_empty((len(positions), config.num_range_bins, len(centers), 2))
    jac = centers.new_empty((*out.shape[:-1], 6, 2))
    cfg = _lib.fmcw(config, fresnel_n)
    _lib.call(
        "dsplat_surfel_jacobian",
        _lib.ctypes.byref(cfg),
        _lib.ptr(positions.contiguous()),
        len(positions),
        _lib.ptr(centers.contiguous()),
        _lib.ptr(normals.contiguous()),
        len(centers),
        _lib.ptr(out),
        _lib.ptr(jac),
        _lib.stream(centers.device),
    )
    jac = torch.view_as_complex(jac)
    return torch.view_as_complex(out), jac[..., :3], jac[..., 3:]


def _workspace(like, apertures, count, width):
    size = _lib.library().dsplat_surfel_correlate_workspace(apertures, count, width)
    return like.new_empty(max(1, size // 4)), size


def correlate(config, positions, centers, normals, residual, fresnel_n):
    """conj(column) . residual and |column|^2 per candidate, without area (kernels/surfel.py:242-260)."""
    _check(positions, centers, normals)
    count = len(centers)
    output = centers.new_empty((count, 3))
    workspace, size = _workspace(centers, len(positions), count, 3)
    cfg = _lib.fmcw(config, fresnel_n)
    _lib.call(
        "dsplat_surfel_correlate",
        _lib.ctypes.byref(cfg),
        _lib.ptr(positions.contiguous()),
        len(positions),
        _lib.ptr(centers.contiguous()),
        _lib.ptr(normals.contiguous()),
        count,
        _lib.ptr(interleaved(residual.reshape(len(positions), config.num_range_bins))),
        _lib.ptr(output),
        _lib.ptr(workspace),
        size,
        _lib.stream(centers.device),
    )
    return torch.complex(output[:, 0], output[:, 1]), output[:, 2]


def correlate_orientations(config, positions, centers, normals, bank, residual, fresnel_n, table=None):
    """Correlations and norms [S, 1 + bank] for the current normal and every bank normal.

    table, an fmcw.RangeTable, replaces the per-bin sum by FFT-interpolated range
    correlations; it only ranks proposals. Without area (kernels/surfel.py:263-290).
    """
    _check(positions, centers, normals)
    count, extra = len(centers), len(bank)
    output = centers.new_empty((count, extra + 1, 3))
    workspace, size = _workspace(centers, len(positions), count, 3 * (extra + 1))
    cfg = _lib.fmcw(config, fresnel_n)
    res = None if table is not None else interleaved(residual.reshape(len(positions), config.num_range_bins))
    _lib.call(
        "dsplat_surfel_correlate_orientations",
        _lib.ctypes.byref(cfg),
        _lib.ptr(positions.contiguous()),
        len(positions),
        _lib.ptr(centers.contiguous()),
        _lib.ptr(normals.contiguous()),
        count,
        _lib.ptr(bank.to(torch.float32).contiguous()) if extra else None,
        extra,
        _lib.ptr(res),
        None if table is None else _lib.ptr(table.correlation.contiguous()),
        None if table is None else _lib.ptr(table.norm.contiguous()),
        1 if table is None else int(table.oversample),
        _lib.ptr(output),
        _lib.ptr(workspace),
        size,
        _lib.stream(centers.device),
    )
    return torch.complex(output[..., 0], output[..., 1]), output[..., 2]


class CertificateOperands:
    """The sample-domain operand tiles of one residual [A, K] (dsplat_certificate_operands),
    built once and shared by every certificate_gradient launch against that residual."""

    def __init__(self, config, residual, fresnel_n=0.0):
        self.cfg = _lib.fmcw(config, fresnel_n)
        self.apertures = residual.shape[0]
        size = _lib.library().dsplat_certificate_operands_bytes(_lib.ctypes.byref(self.cfg), self.apertures)
        if size == 0:
            raise ValueError("the sample-domain certificate kernels do not serve this chirp")
        self.buffer = torch.empty(size, dtype=torch.uint8, device=residual.device)
        _lib.call(
            "dsplat_certificate_operands", _lib.ctypes.byref(self.cfg),
            _lib.ptr(interleaved(residual.reshape(self.apertures, config.num_range_bins))), self.apertures,
            _lib.ptr(self.buffer), size, _lib.stream(residual.device),
        )  # fmt: skip


def certificate_gradient(positions, centers, normals, operands):
    """[C, 21] = (corr, norm, dcorr re / im [6], dnorm [6]) of C candidates against the residual
    of `operands` (CertificateOperands), without area: correlate_gradient's columns with the
    normals, on the sample-domain matrix-core kernel."""
    _check(positions, centers, normals)
    count = len(centers)
    out = centers.new_empty((count, 21))
    if count == 0:
        return out
    cfg = operands.cfg
    size = _lib.library().dsplat_certificate_gradient_workspace(_lib.ctypes.byref(cfg), len(positions), count)
    workspace = centers.new_empty(max(1, size // 4))
    _lib.call(
        "dsplat_certificate_gradient", _lib.ctypes.byref(cfg), _lib.ptr(positions.contiguous()), len(positions),
        _lib.ptr(centers.contiguous()), _lib.ptr(normals.contiguous()), count, _lib.ptr(operands.buffer),
        _lib.ptr(out), _lib.ptr(workspace), size, _lib.stream(centers.device),
    )  # fmt: skip
    return out


def sample_domain(config, fresnel_n=0.0):
    """Whether the sample-domain kernels serve this chirp (its splat form and ADC samples)."""
    cfg = _lib.fmcw(config, fresnel_n)
    return _lib.library().dsplat_project_form(_lib.ctypes.byref(cfg), 2) == 0


def sample_basis(config, basis, fresnel_n=0.0):
    """Sample-domain form [rows, 2 * adc_samples] of an interleaved real row basis [rows, 2K]
    (B = U^H F, F[k][n] = e^{-2 pi i n k / N}, float64 sums) for project_design's `sampled`
    argument, or None when only the bin-domain projection serves this chirp."""
    rows = basis.shape[0]
    cfg = _lib.fmcw(config, fresnel_n)
    if _lib.library().dsplat_project_form(_lib.ctypes.byref(cfg), rows) != 0:
        return None
    out = torch.empty((rows, 2 * int(config.adc_samples)), dtype=torch.float32, device=basis.device)
    _lib.call(
        "dsplat_sample_basis", _lib.ctypes.byref(cfg), _lib.ptr(basis.float().contiguous()), rows, _lib.ptr(out),
        _lib.stream(basis.device),
    )  # fmt: skip
    return out


def project_design(config, positions, centers, normals, areas, basis, fresnel_n, sampled=None, column_max=None):
    """Row-projected real design [rows * A, S] (row j * A + a) of a basis [rows, 2K] in its
    interleaved real form, atoms generated on chip (fmcw.py:243-261 blocks + reduce_major).
    With `sampled` (sample_basis of the same basis) the sample-domain kernel runs instead,
    and then `column_max` (zeroed int32 [S]) receives the float bits of max |out[:, s]|."""
    _check(positions, centers, normals)
    rows = basis.shape[0]
    out = centers.new_empty((rows * len(positions), len(centers)))
    cfg = _lib.fmcw(config, fresnel_n)
    head = (
        _lib.ctypes.byref(cfg), _lib.ptr(positions.contiguous()), len(positions), _lib.ptr(centers.contiguous()),
        _lib.ptr(normals.contiguous()), _lib.ptr(areas.float().contiguous()), len(centers),
    )  # fmt: skip
    if sampled is None:
        if column_max is not None:
            raise ValueError("column maxima come with the sample-domain projection only")
        _lib.call("dsplat_project_design", *head, _lib.ptr(basis.float().contiguous()), rows, _lib.ptr(out),
                  _lib.stream(centers.device))  # fmt: skip
    else:
        if column_max is not None and (column_max.dtype != torch.int32 or column_max.shape != (len(centers),)):
            raise ValueError("column_max must be an int32 [S] tensor")
        _lib.call("dsplat_project_sampled", *head, _lib.ptr(sampled.float().contiguous()), rows, _lib.ptr(out),
                  _lib.ptr(column_max), _lib.stream(centers.device))  # fmt: skip
    return out


def project_split_rows(rows):
    """Rows per aperture of a split block: rows rounded up to 32."""
    return int(_lib.library().dsplat_project_split_rows(rows))


def project_split(config, positions, centers, normals, areas, sampled, rows, fresnel_n, format="bf16", exponents=None,
                  out=None, col0=0):
    """The sample-domain projected block of project_design in the operand form of
    blas.syrk_split: a SplitBlock of m = A * rows_pad rows (rows_pad = rows rounded up to 32;
    projected row j of aperture a is row a * rows_pad + j, padding rows zero), S columns, in
    BF16, or in FP16 with every column scaled by 2^exponents[s] (blas.projection_exponents).

    out = (hi, lo), existing [m / 32, ld, 32] arrays, takes the S columns at columns
    col0 .. col0 + S - 1 instead (ld >= col0 + S; the other columns are left as they are);
    nothing is returned then."""
    from .blas import SPLIT_FORMATS, SplitBlock

    _check(positions, centers, normals)
    code, dtype = SPLIT_FORMATS[format]
    cfg = _lib.fmcw(config, fresnel_n)
    m = project_split_rows(rows) * len(positions)
    if out is None:
        hi = torch.empty((m // 32, len(centers), 32), dtype=dtype, device=centers.device)
        lo = torch.empty_like(hi)
        first_hi, first_lo = _lib.ptr(hi), _lib.ptr(lo)
    else:
        hi, lo = out
        if hi.dtype != dtype or lo.dtype != dtype or hi.shape != lo.shape or hi.shape[0] != m // 32:
            raise ValueError(f"project_split: out is two {dtype} [{m // 32}, ld, 32] arrays")
        if col0 < 0 or col0 + len(centers) > hi.shape[1]:
            raise ValueError("project_split: columns col0 .. col0 + S - 1 must lie within ld")
        step = 32 * hi.element_size() * col0  # column col0 of tile 0; ld stays the tiles' stride
        first_hi = _lib.ctypes.c_void_p(_lib.ptr(hi).value + step)
        first_lo = _lib.ctypes.c_void_p(_lib.ptr(lo).value + step)
    if exponents is not None:
        exponents = exponents.float().contiguous()
    _lib.call(
        "dsplat_project_sampled_split", _lib.ctypes.byref(cfg), _lib.ptr(positions.contiguous()), len(positions),
        _lib.ptr(centers.contiguous()), _lib.ptr(normals.contiguous()), _lib.ptr(areas.float().contiguous()),
        len(centers), _lib.ptr(sampled.float().contiguous()), rows, code, _lib.ptr(exponents), first_hi, first_lo,
        hi.shape[1], _lib.stream(centers.device),
    )  # fmt: skip
    if out is None:
        return SplitBlock(hi, lo, m, format=format, exponents=exponents)
    return None


def orient_select(config, positions, centers, normals, areas, bank, residual, fresnel_n, table=None, real_part=False):
    """Best orientation per candidate in one launch: (value [S] float64, index [S] int64),
    value = |s|^2 / max(area |a|^2, 1e-60) with s = sqrt(area) a^H r (or its real part),
    index 0 for the own normal and 1 + b for bank normal b (certificate.py:73-103)."""
    _check(positions, centers, normals)
    count, extra = len(centers), len(bank)
    value = torch.empty(count, dtype=torch.float64, device=centers.device)
    index = torch.empty(count, dtype=torch.int32, device=centers.device)
    if count == 0:
        return value, index.long()
    cfg = _lib.fmcw(config, fresnel_n)
    res = None if table is not None else interleaved(residual.reshape(len(positions), config.num_range_bins))
    _lib.call(
        "dsplat_surfel_orient_select", _lib.ctypes.byref(cfg), _lib.ptr(positions.contiguous()), len(positions),
        _lib.ptr(centers.contiguous()), _lib.ptr(normals.contiguous()), _lib.ptr(areas.float().contiguous()), count,
        _lib.ptr(bank.to(torch.float32).contiguous()) if extra else None, extra, _lib.ptr(res),
        None if table is None else _lib.ptr(table.correlation.contiguous()),
        None if table is None else _lib.ptr(table.norm.contiguous()),
        1 if table is None else int(table.oversample), int(bool(real_part)), _lib.ptr(value), _lib.ptr(index),
        _lib.stream(centers.device),
    )  # fmt: skip
    return value, index.long()


def correlate_gradient(config, positions, centers, normals, residual, fresnel_n, include_normals):
    """Correlation, norm and their center (and normal) derivatives [S, 3 or 6], without area
    (kernels/surfel.py:293-315)."""
    _check(positions, centers, normals)
    width = 6 if include_normals else 3
    count = len(centers)
    output = centers.new_empty((count, 3 + 3 * width))
    workspace, size = _workspace(centers, len(positions), count, 3 + 3 * width)
    cfg = _lib.fmcw(config, fresnel_n)
    _lib.call(
        "dsplat_surfel_correlate_gradient",
        _lib.ctypes.byref(cfg),
        _lib.ptr(positions.contiguous()),
        len(positions),
        _lib.ptr(centers.contiguous()),
        _lib.ptr(normals.contiguous()),
        count,
        _lib.ptr(interleaved(residual.reshape(len(positions), config.num_range_bins))),
        1 if include_normals else 0,
        _lib.ptr(output),
        _lib.ptr(workspace),
        size,
        _lib.stream(centers.device),
    )
    return (
        torch.complex(output[:, 0], output[:, 1]),
        output[:, 2],
        torch.complex(output[:, 3 : 3 + width], output[:, 3 + width : 3 + 2 * width]),
        output[:, 3 + 2 * width :],
    )


SCAN_CHUNKS = (512, 2048, 8192, 32768)  # autotuned: candidates per workgroup (bin domain) or apertures per range (sample domain)


def scan_batched(config, scans, centers, normals, fresnel_n, chunk=None):
    """corr [n, C] complex and norm [n, C] of every candidate against each of n scans
    (positions [A_i, 3], residual [A_i, K]) in one launch over uneven aperture counts,
    without area. chunk None autotunes the candidates per workgroup."""
    out = scan_rows(config, scans, centers, normals, fresnel_n, chunk)
    return torch.complex(out[..., 0], out[..., 1]), out[..., 2]


def scan_rows(config, scans, centers, normals, fresnel_n, chunk=None):
    """The launch of scan_batched: rows [n, C, 3] float32 = (corr.re, corr.im, norm)."""
    import ctypes

    from .. import autotune

    _check(scans[0][0], centers, normals)
    count, n = len(centers), len(scans)
    positions = [p.float().contiguous() for p, _ in scans]
    residuals = [interleaved(r.reshape(len(p), config.num_range_bins)) for p, r in scans]
    pos_ptrs = (ctypes.c_void_p * n)(*[p.data_ptr() for p in positions])
    res_ptrs = (ctypes.c_void_p * n)(*[r.data_ptr() for r in residuals])
    apertures = (ctypes.c_int32 * n)(*[len(p) for p in positions])
    cfg = _lib.fmcw(config, fresnel_n)
    size = _lib.library().dsplat_scan_workspace(_lib.ctypes.byref(cfg), n, apertures, count)
    workspace = centers.new_empty(max(1, size // 4))
    out = centers.new_empty((n, count, 3))
    centers, normals = centers.contiguous(), normals.contiguous()

    def run(value):
        _lib.call(
            "dsplat_scan_batched", _lib.ctypes.byref(cfg), n, pos_ptrs, res_ptrs, apertures, _lib.ptr(centers),
            _lib.ptr(normals), count, int(value), _lib.ptr(out), _lib.ptr(workspace), size,
            _lib.stream(centers.device),
        )  # fmt: skip

    if chunk is None:
        # the form and the sample count pick the kernel, and with it what chunk means
        key = (f"C{count}|A{'-'.join(str(len(p)) for p in positions)}|K{config.num_range_bins}|P{config.pad_factor}"
               f"|S{config.adc_samples}|F{cfg.splat_form}")
        chunk = autotune.tune("scan_batched", key, list(SCAN_CHUNKS), run, centers.device)
    run(chunk)
    return out
