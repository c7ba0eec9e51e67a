"""Dense products on MFMA: the symmetric rank-k update the reference binds from cuBLAS
(dsplat/kernels/cublas.py) and the fused split-precision Gram."""

import torch

from .. import _lib


SYRK_MODES = {"fp32": 0, "bf16x3": 1, "fp16x3": 2}


def _column_max(column_max, n):
    if column_max.dtype != torch.int32 or column_max.shape != (n,) or not column_max.is_contiguous():
        raise ValueError("column_max must be a contiguous int32 [n] tensor of float bits")
    return _lib.ptr(column_max)


def syrk(rows, out, mode="fp32", column_max=None):
    """Add rows^T rows to the upper triangle of out [n, n]; the lower triangle is untouched
    (kernels/cublas.py:31-44). rows [m, n] float32, out float32 contiguous. mode: fp32
    (FP32 MFMA), bf16x3 or fp16x3 (hi + lo split, three MFMA products, FP32 accumulation).
    column_max: the float bits of max |rows[:, c]| per column (int32 [n], as the sample-domain
    projection writes them), so the FP16 mode skips its own pass over rows."""
    if rows.dtype != torch.float32 or out.dtype != torch.float32:
        raise TypeError("syrk takes float32 tensors")
    if not out.is_contiguous() or out.shape != (rows.shape[1], rows.shape[1]):
        raise ValueError("out must be a contiguous [n, n] tensor")
    code = SYRK_MODES[mode]
    rows = rows.contiguous()
    m, n = rows.shape
    size = _lib.library().dsplat_syrk_workspace(m, n, code)
    workspace = torch.empty(max(1, size // 4), dtype=torch.float32, device=rows.device)
    tail = (_lib.ptr(out), _lib.ptr(workspace), size, _lib.stream(rows.device))
    if column_max is None:
        _lib.call("dsplat_syrk", _lib.ptr(rows), m, n, code, *tail)
    else:
        _lib.call("dsplat_syrk_scaled", _lib.ptr(rows), m, n, code, _column_max(column_max, n), *tail)
    return out


def gemm_tn(rows, col0, out, mode="fp32", column_max=None):
    """out [n, n - col0] += rows^T rows[:, col0:] on the SYRK tiles (rows [m, n] float32);
    column_max as for syrk."""
    if rows.dtype != torch.float32 or out.dtype != torch.float32:
        raise TypeError("gemm_tn takes float32 tensors")
    m, n = rows.shape
    if not out.is_contiguous() or out.shape != (n, n - col0):
        raise ValueError("out must be a contiguous [n, n - col0] tensor")
    code = SYRK_MODES[mode]
    rows = rows.contiguous()
    size = _lib.library().dsplat_gemm_tn_workspace(m, n, col0, code)
    workspace = torch.empty(max(1, size // 4), dtype=torch.float32, device=rows.device)
    tail = (_lib.ptr(out), _lib.ptr(workspace), size, _lib.stream(rows.device))
    if column_max is None:
        _lib.call("dsplat_gemm_tn", _lib.ptr(rows), m, n, int(col0), code, *tail)
    else:
        _lib.call("dsplat_gemm_tn_scaled", _lib.ptr(rows), m, n, int(col0), code, _column_max(column_max, n), *tail)
    return out


SPLIT_FORMATS = {"bf16": (0, torch.bfloat16), "fp16": (1, torch.float16)}


class SplitBlock:
    """A block P [m, n] in the operand form of syrk_split: 16-bit hi and lo [m / 32, ld, 32]
    (32-row column tiles: element (k, c) at [k // 32, c, k % 32]; ld >= n) with hi + lo ~ P in
    BF16, or hi + lo ~ P times 2^exponents[c] per column in FP16 (exponents float [n])."""

    def __init__(self, hi, lo, m, columns=None, format="bf16", exponents=None):
        code, dtype = SPLIT_FORMATS[format]
        if hi.dtype != dtype or lo.dtype != dtype or hi.shape != lo.shape:
            raise TypeError(f"a {format} split block is two {dtype} [m / 32, ld, 32] tensors")
        if not (hi.is_contiguous() and lo.is_contiguous()) or hi.ndim != 3 or hi.shape[2] != 32:
            raise ValueError("split block: contiguous [m / 32, ld, 32] tensors")
        if m % 32 or hi.shape[0] != m // 32:
            raise ValueError("split block: m is a multiple of 32 and hi has m / 32 tiles")
        if format == "fp16" and (exponents is None or exponents.dtype != torch.float32):
            raise ValueError("an fp16 split block carries float32 column exponents")
        self.hi, self.lo, self.m = hi, lo, int(m)
        self.columns = hi.shape[1] if columns is None else int(columns)
        self.format, self.code, self.exponents = format, code, exponents

    @staticmethod
    def of(rows, format="bf16", exponents=None):
        """The split form of a float32 [m, n] matrix (m a multiple of 32)."""
        m, n = rows.shape
        t = rows.float().reshape(m // 32, 32, n).transpose(1, 2).contiguous()
        if exponents is not None:
            t = t * torch.exp2(exponents)[None, :, None]
        dtype = SPLIT_FORMATS[format][1]
        hi = t.to(dtype)
        lo = (t - hi.float()).to(dtype)
        return SplitBlock(hi, lo, m, format=format, exponents=exponents)

    def dense(self):
        """float32 [m, n] of hi + lo (unscaled)."""
        t = self.hi[:, : self.columns].float() + self.lo[:, : self.columns].float()
        if self.exponents is not None:
            t = t * torch.exp2(-self.exponents[: self.columns])[None, :, None]
        return t.transpose(1, 2).reshape(self.m, self.columns)


def syrk_split(block, out, col0=None):
    """Upper triangle of out [n, n] += P^T P (col0 None), or out [n, n - col0] +=
    P^T P[:, col0:], for a SplitBlock P: hi hi + hi lo + lo hi on the 16-bit matrix cores with
    FP32 accumulation (kernels/cublas.py:31-44 SSYRK; replacement.py:178-199 cross products)."""
    n = block.columns
    tail = n if col0 is None else n - col0
    if out.dtype != torch.float32 or not out.is_contiguous() or out.shape != (n, tail):
        raise ValueError("out must be a contiguous float32 [n, n - col0] tensor")
    c0 = -1 if col0 is None else int(col0)
    size = _lib.library().dsplat_syrk_split_workspace(block.m, n, c0)
    workspace = torch.empty(max(1, size // 4), dtype=torch.float32, device=out.device)
    _lib.call("dsplat_syrk_split", block.code, _lib.ptr(block.hi), _lib.ptr(block.lo), block.hi.shape[1], block.m, n,
              c0, _lib.ptr(block.exponents), _lib.ptr(out), _lib.ptr(workspace), size,
              _lib.stream(out.device))  # fmt: skip
    return out


def projection_exponents(config, positions, centers, areas, target=2.0**14):
    """Power-of-two column scales for FP16 split blocks of the row-projected design: 2^e_s
    times any projected entry of column s stays below target. A projected entry is the
    weight sqrt(area) amp rot of its pair times U^H d(k0), where |amp| <= lambda / (8 pi r)
    (cos^2 and the Fresnel factor are at most 1) and |U^H d| <= |d|_2 <= sqrt(pad) (U has
    orthonormal columns; the unit-peak Dirichlet response has energy at most n_fft /
    adc_samples over the bins); r is at least the distance from the centre to the bounding
    box of the apertures."""
    lower, upper = positions.min(0).values.double(), positions.max(0).values.double()
    c = centers.double()
    nearest = torch.minimum(torch.maximum(c, lower), upper)
    r_min = (c - nearest).norm(dim=1).clamp(min=1e-9)
    bound = areas.double().sqrt() * config.lambda_m / (8 * torch.pi * r_min) * config.pad_factor**0.5
    return torch.floor(torch.log2(target / bound.clamp(min=1e-300))).clamp(-120, 120).float()


GRAM_MODES = {"fp32": 0, "bf16x3": 1, "fp16x3": 2}
FP16_TARGET = 2.0**14  # largest scaled atom magnitude in FP16 mode (max finite 65504)


def column_exponents(config, positions, centers, areas):
    """Power-of-two column scales for the FP16 split: 2^e_s |atom| <= 2^14 for every atom.

    |D| <= 1, cos^2 F <= 1, so |atom| <= sqrt(area) lambda / (8 pi r_min) with r_min the
    column's nearest aperture.
    """
    r_min = torch.cdist(centers.double(), positions.double()).min(1).values.clamp(min=1e-9)
    bound = areas.double().sqrt() * config.lambda_m / (8 * torch.pi * r_min)
    return torch.floor(torch.log2(FP16_TARGET / bound.clamp(min=1e-300))).clamp(-120, 120).float()


def gram(config, positions, centers, normals, areas, mode="bf16x3", real_only=False, scale=1.0, fresnel_n=0.0):
    """scale * A^H A of the surfel dictionary (atoms include sqrt(area)), atoms generated
    on chip and contracted on MFMA; complex [S, S] Hermitian, or real [S, S] = Re when
    real_only. mode: fp32 | bf16x3 | fp16x3 (hi+lo split, FP32 accumulation)."""
    count = len(centers)
    code = GRAM_MODES[mode]
    positions, centers, normals = (t.float().contiguous() for t in (positions, centers, normals))
    areas = areas.float().contiguous()
    out = centers.new_empty((count, count) if real_only else (count, count, 2))
    size = _lib.library().dsplat_gram_workspace(len(positions), count)
    workspace = centers.new_empty(max(1, size // 4))
    exponents = column_exponents(config, positions, centers, areas) if code == 2 else None
    cfg = _lib.fmcw(config, fresnel_n)
    _lib.call(
        "dsplat_gram", _lib.ctypes.byref(cfg), _lib.ptr(positions), len(positions), _lib.ptr(centers),
        _lib.ptr(normals), _lib.ptr(areas), count, code, 1 if real_only else 0, float(scale),
        _lib.ptr(exponents), _lib.ptr(out), _lib.ptr(workspace), size, _lib.stream(centers.device),
    )  # fmt: skip
    return out if real_only else torch.view_as_complex(out)
