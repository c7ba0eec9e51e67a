"""FMCW sensing: the chirp, a scan, and the surfel measurement model the solver works with.

Public surface of the reference's dsplat/fmcw.py:29-355 (FMCWConfig, Scan, FMCWModel and
the row basis of the range response). A linear chirp of bandwidth B around fc is sampled by
adc_samples ADC samples and zero-padded to n_fft; a scan holds the positive range bins of
every aperture, data[A, K] = FFT(ADC, n_fft)[:K] / adc_samples, and a reflector at range r
lands at bin r / spacing with spacing = c / (2 B pad_factor).

FMCWModel is organized in three parts: the splat products (render, JVP, VJP, dictionary
rows and Jacobians), the least-squares stream (row-projected blocks on the range subspace,
in float32 or in the split BF16 operand form of the MFMA normal equations), and the search
support (correlations, one-launch orientation choice and quasi-Newton ascent, the range
gate and the FFT range table). Every kernel runs through the C ABI on torch's stream.
"""

import math
from dataclasses import dataclass, field, replace
from math import ceil, floor
from typing import NamedTuple

import torch
import torch.nn.functional as F

from . import surfel
from .kernels import search
from .kernels.response import range_response


@dataclass(frozen=True)
class FMCWConfig:
    """Chirp constants (fmcw.py:29-65 of the reference)."""

    fc_hz: float = 120e9
    bandwidth_hz: float = 15e9
    adc_samples: int = 128
    sweep_time_s: float = 1e-5
    adc_start_time_s: float = 0.0
    pad_factor: int = 2
    c0_m_s: float = 299792458.0
    # Kernel family of every splat, correlation and projection call of this chirp (not in
    # the reference, which has one): "samples" evaluates the Dirichlet kernel as its sample
    # sum on the matrix cores where the chirp has 128 or 256 ADC samples (other chirps use
    # "bins"), "bins" the bin-domain cotangent form. Both agree with float64 to ~1e-7.
    splat_form: str = "samples"

    @property
    def lambda_m(self):
        return self.c0_m_s / self.fc_hz

    @property
    def slope_hz_s(self):
        return self.bandwidth_hz / self.sweep_time_s

    @property
    def sample_rate_hz(self):
        return self.adc_samples / self.sweep_time_s

    @property
    def n_fft(self):
        return self.adc_samples * self.pad_factor

    @property
    def num_range_bins(self):
        return self.n_fft // 2

    @property
    def range_bin_spacing_m(self):
        return self.c0_m_s / (2 * self.bandwidth_hz * self.pad_factor)

    @property
    def range_resolution_m(self):
        return self.c0_m_s / (2 * self.bandwidth_hz)


@dataclass(frozen=True)
class Scan:
    """Range profiles data [A, K] measured at scan positions [A, 3] (metres)."""

    data: torch.Tensor
    scan_positions_m: torch.Tensor
    fmcw_config: FMCWConfig


# --- The range subspace ---------------------------------------------------------------------

# Samples of the peak position per range bin in the response matrix whose SVD gives the basis.
PEAKS_PER_BIN = 16


class RangeSubspace:
    """An orthonormal basis U [K, r] of the range responses of every peak position in a band
    of bins. Its adjoint projects a block's observations (whole apertures, K rows each) onto
    r rows per aperture without changing the normal equations to float32 precision, so the
    least-squares stream shrinks by K / r (fmcw.py:77-118 of the reference).

    adjoint      complex [r, K], U^H
    interleaved  real [2 r, 2 K], U^H acting on (re, im) row pairs
    """

    def __init__(self, adjoint, interleaved):
        self.adjoint = adjoint
        self.interleaved = interleaved
        self._sampled = {}

    @property
    def rank(self):
        return self.adjoint.shape[0]

    def reduce(self, values):
        """Complex rows [apertures * K, ...] to [apertures * r, ...]."""
        tail = values.shape[1:]
        per_aperture = values.reshape(-1, self.adjoint.shape[1], tail.numel())
        return (self.adjoint @ per_aperture).reshape(-1, *tail)

    def reduce_major(self, rows):
        """Real rows ordered bin by bin, row (2 k + part) * apertures + a, in one product."""
        tail = rows.shape[1:]
        return (self.interleaved @ rows.reshape(self.interleaved.shape[1], -1)).reshape(-1, *tail)

    def sampled(self, config, fresnel_n):
        """The sample-domain form of the basis for the projection kernels (one conversion per
        chirp), or None where only the bin-domain projection serves the chirp."""
        key = (config.adc_samples, config.pad_factor, config.splat_form)
        if key not in self._sampled:
            self._sampled[key] = surfel.kernel.sample_basis(config, self.interleaved, fresnel_n)
        return self._sampled[key]


def response_basis(config, low, high, device):
    """The range subspace of peak positions in [low, high] bins, or None if it spans every bin.

    Float64 responses come from the range_response kernel and their SVD runs on the host in
    float64; singular values above eps32 / 128 of the largest are kept (fmcw.py:101-118)."""
    count = max(2 * config.adc_samples, int(PEAKS_PER_BIN * (high - low)) + 1)
    responses = range_response(config, low, high, count, device).cpu()
    left, singular, _ = torch.linalg.svd(responses, full_matrices=False)
    rank = int((singular > torch.finfo(torch.float32).eps / 128 * singular[0]).sum())
    if rank >= config.num_range_bins:
        return None
    adjoint = left[:, :rank].mH
    # (re, im) row pairs: [[Re, -Im], [Im, Re]] blocks of U^H
    real = torch.zeros((2 * rank, 2 * config.num_range_bins), dtype=torch.float64)
    for part, (even, odd) in enumerate(((adjoint.real, -adjoint.imag), (adjoint.imag, adjoint.real))):
        real[part::2, 0::2], real[part::2, 1::2] = even, odd
    return RangeSubspace(adjoint.to(device, torch.complex64), real.to(device, torch.float32))


def _band(positions, lower, upper, spacing, n_fft):
    """Per aperture, the nearest (minus one bin) and farthest (plus one bin) range bin of the
    box [lower, upper]; None when the box's band covers the whole FFT."""
    lo, hi = positions.new_tensor(lower).expand(3), positions.new_tensor(upper).expand(3)
    closest = torch.minimum(torch.maximum(positions, lo), hi)
    corner = torch.where((positions - lo).abs() > (positions - hi).abs(), lo, hi)
    near = ((positions - closest).norm(dim=1) / spacing - 1).clamp(min=0)
    far = (positions - corner).norm(dim=1) / spacing + 1
    if float(far.max()) - float(near.min()) >= n_fft:
        return None
    return near, far


class RangeTable(NamedTuple):
    """FFT-interpolated range correlations of a residual; they only rank proposals."""

    correlation: torch.Tensor  # [apertures, n_fft * oversample, 2]
    norm: torch.Tensor  # [n_fft * oversample]; keeps the observed-bin mask
    oversample: int


# --- The measurement model ------------------------------------------------------------------


class _SplatProducts:
    """render / jvp / vjp and dictionary rows of whole or sampled apertures."""

    def render(self, params, coefficients):
        return surfel.render(*self.sensor, params, coefficients, self.fresnel_n)

    def jvp(self, params, coefficients, tangent, delta_coefficients):
        return surfel.jvp(*self.sensor, params, coefficients, tangent, delta_coefficients, self.fresnel_n)

    def vjp(self, params, coefficients, upstream):
        return surfel.vjp(*self.sensor, params, coefficients, upstream, self.fresnel_n)

    def _aperture_rows(self, rows):
        """(positions of the apertures rows touch, row -> aperture, row -> bin)."""
        bins = self.shape[1]
        apertures, owner = torch.unique(rows // bins, return_inverse=True)
        return self.scan.scan_positions_m[apertures], owner, rows % bins

    def design(self, params, rows):
        """Complex dictionary rows [rows, S]; a slice must cover whole apertures."""
        config, positions = self.sensor
        bins = self.shape[1]
        if isinstance(rows, slice):
            block = positions[rows.start // bins : -(-rows.stop // bins)]
            return surfel.design(config, block, params, self.fresnel_n)
        block, owner, bin_of = self._aperture_rows(rows)
        columns = surfel.design(config, block, params, self.fresnel_n)
        return columns.reshape(len(block), bins, -1)[owner, bin_of]

    def jacobian(self, params, rows, atoms, keys):
        block, owner, bin_of = self._aperture_rows(rows)
        chosen = {key: value[atoms] for key, value in params.items()}
        derivatives = surfel.jacobian(self.scan.fmcw_config, block, chosen, self.fresnel_n)
        return {key: derivatives[key][owner, bin_of] for key in keys}


class _LeastSquaresStream:
    """Row-projected least-squares blocks of whole apertures (fmcw.py:196-261)."""

    def prepare(self, specs):
        """This model with the range subspace of every centre the centres' box allows (the band
        of its nearest and farthest point from any aperture, one bin of margin)."""
        lower, upper = specs["centers"].lower, specs["centers"].upper
        config, positions = self.sensor
        low, high, bands = 0.0, float(config.n_fft), None
        if lower is not None and upper is not None:
            band = _band(positions, lower, upper, config.range_bin_spacing_m, config.n_fft)
            if band is not None:
                low, high = float(band[0].min()), float(band[1].max())
                bands = (band[0].cpu().numpy(), band[1].cpu().numpy())
        return replace(self, basis=response_basis(config, low, high, positions.device), bands=bands, block_bases={},
                       split_cache=SplitCache())

    def block_basis(self, rows):
        """The range subspace of a block of whole apertures: the band of its own apertures,
        rounded out to pad-factor steps, where that is smaller than the shared one."""
        if self.bands is None or self.basis is None:
            return self.basis
        bins, step = self.shape[1], self.scan.fmcw_config.pad_factor
        first, last = rows.start // bins, -(-rows.stop // bins)
        key = (
            floor(self.bands[0][first:last].min() / step) * step,
            ceil(self.bands[1][first:last].max() / step) * step,
        )
        if key not in self.block_bases:
            config, positions = self.sensor
            own = response_basis(config, float(key[0]), float(key[1]), positions.device)
            self.block_bases[key] = own if own is not None and own.rank < self.basis.rank else self.basis
        return self.block_bases[key]

    def blocks(self, params, target, real, chunk):
        """(design rows, observations) of `chunk` rows at a time, projected on the range
        subspace; real blocks are ordered bin by bin so that one product projects them."""
        config, positions = self.sensor
        bins = self.shape[1]
        flat = target.reshape(-1)
        for first in range(0, self.size, chunk):
            rows = slice(first, min(first + chunk, self.size))
            basis = self.block_basis(rows)
            if real:
                block = positions[rows.start // bins : rows.stop // bins]
                a = surfel.design(config, block, params, self.fresnel_n, real=True)
                y = torch.view_as_real(flat[rows]).reshape(len(block), 2 * bins).T.reshape(-1)
                yield (a, y) if basis is None else (basis.reduce_major(a), basis.reduce_major(y))
            else:
                a, y = self.design(params, rows), flat[rows]
                yield (a, y) if basis is None else (basis.reduce(a), basis.reduce(y))

    def _aperture_steps(self, chunk):
        return range(0, self.apertures, max(1, chunk // self.shape[1]))

    def projected_blocks(self, params, chunk, column_max=False):
        """The real row-projected design blocks of `blocks` (without y) on the shared
        subspace, atoms generated on chip and never written; None without a subspace. With
        column_max, (block, maxima) pairs, maxima the float bits of each column's largest
        |entry| (int32 [S]) where the sample-domain projection forms them, else None."""
        if self.basis is None:
            return None
        config, positions = self.sensor
        step = max(1, chunk // self.shape[1])
        sampled = self.basis.sampled(config, self.fresnel_n)
        count = len(params["centers"])

        def block(first):
            maxima = torch.zeros(count, dtype=torch.int32, device=positions.device) if (
                column_max and sampled is not None) else None  # fmt: skip
            a = surfel.kernel.project_design(
                config, positions[first : first + step], params["centers"], params["normals"], params["areas"],
                self.basis.interleaved, self.fresnel_n, sampled, maxima,
            )  # fmt: skip
            return (a, maxima) if column_max else a

        return (block(first) for first in self._aperture_steps(chunk))

    def split_blocks(self, params, chunk, format="bf16", known=None):
        """The same blocks in the operand form of the split SYRK (kernels.blas.SplitBlock:
        16-bit hi / lo in 32-row column tiles, written by the projection kernel itself; FP16
        with column scales that hold for every aperture of the model), or None where the
        sample-domain projection does not serve the chirp.

        A model with a split_cache projects every aperture into it at once and returns views of
        it: columns the cache already holds for `known` (the leading atoms of params, e.g. the
        support before pooled candidates) are not projected again. The values are the same
        either way: every column is projected on its own."""
        if self.basis is None:
            return None
        config, positions = self.sensor
        sampled = self.basis.sampled(config, self.fresnel_n)
        if sampled is None:
            return None
        from .kernels.blas import SplitBlock, projection_exponents

        step = max(1, chunk // self.shape[1])
        rows = 2 * self.basis.rank
        centers, normals, areas = params["centers"], params["normals"], params["areas"]

        def exponents_of(first=0):
            if format != "fp16":
                return None
            return projection_exponents(config, positions, centers[first:], areas[first:])

        cache = self.split_cache
        if cache is None:
            return (
                surfel.kernel.project_split(config, positions[first : first + step], centers, normals, areas, sampled,
                                            rows, self.fresnel_n, format, exponents_of())
                for first in self._aperture_steps(chunk)
            )  # fmt: skip
        count = len(centers)
        pad = surfel.kernel.project_split_rows(rows)
        first = cache.reserve(self.apertures * pad // 32, count, format, positions.device, known)
        if first < count:
            part = exponents_of(first)
            surfel.kernel.project_split(config, positions, centers[first:], normals[first:], areas[first:], sampled,
                                        rows, self.fresnel_n, format, part, out=(cache.hi, cache.lo), col0=first)
            if part is not None:
                cache.exponents[first:count] = part
        cache.holding(params, count, known)
        exponents = cache.exponents[:count] if format == "fp16" else None
        tiles = step * pad // 32
        return (
            SplitBlock(cache.hi[t : t + tiles], cache.lo[t : t + tiles], min(tiles, len(cache.hi) - t) * 32,
                       columns=count, format=format, exponents=exponents)
            for t in range(0, len(cache.hi), tiles)
        )  # fmt: skip


class SplitCache:
    """The split SYRK operands of the whole row-projected design (every aperture of one model),
    kept in HBM between the normal-equation passes of a fit: hi / lo [A rows_pad / 32, ld, 32]
    and the float32 column exponents of the FP16 form. At 16384 apertures x 64 rows x 2304
    columns that is 9.7 GB.

    `held` lists (atoms, count): columns 0 .. count - 1 are the projections of those atoms
    (centres, normals and areas, matched by tensor identity and in-place version, the tensors
    kept alive here so an identity is never reused). Pooled statistics then project only the
    pooled candidates, and a merged support adopts the columns of the candidates it took."""

    KEYS = ("centers", "normals", "areas")

    def __init__(self):
        self.hi = self.lo = self.exponents = None
        self.format = None
        self.held = []

    @classmethod
    def _token(cls, params):
        return tuple((params[key], params[key]._version) for key in cls.KEYS)

    def _holds(self, params, count):
        token = self._token(params)
        return any(n >= count and all(a is b and u == v for (a, u), (b, v) in zip(t, token)) for t, n in self.held)

    def reserve(self, tiles, count, format, device, known=None):
        """Room for count columns; returns the first column to project (the columns of
        `known` are kept when held)."""
        from .kernels.blas import SPLIT_FORMATS

        dtype = SPLIT_FORMATS[format][1]
        if self.hi is None or self.hi.shape[0] != tiles or self.hi.shape[1] < count or self.format != format:
            ld = -(-count // 256) * 256
            self.hi = self.lo = None
            self.hi = torch.empty((tiles, ld, 32), dtype=dtype, device=device)
            self.lo = torch.empty_like(self.hi)
            self.exponents = torch.zeros(ld, dtype=torch.float32, device=device)
            self.format, self.held = format, []
        if known is not None and self._holds(known, len(known["centers"])):
            return len(known["centers"])
        return 0

    def holding(self, params, count, known=None):
        """Columns 0 .. count - 1 now hold params (and their leading columns `known`)."""
        self.held = [(self._token(params), count)]
        if known is not None:
            self.held.append((self._token(known), len(known["centers"])))

    def adopt(self, pool, params, slots, sources):
        """params took pool column sources[i] into slot slots[i] (the other slots keep the
        pool's columns): move those columns, if the cache holds the pool."""
        count = len(params["centers"])
        if not self._holds(pool, len(pool["centers"])):
            return
        if len(slots):
            for part in (self.hi, self.lo):
                part.index_copy_(1, slots, part.index_select(1, sources))
            self.exponents[slots] = self.exponents[sources]
        self.held = [(self._token(params), count)]


class _SearchSupport:
    """Correlations, batched scoring and the proposal filters of the spatial search."""

    # atoms from which A^H r runs faster as the sample-domain VJP than as the correlation, and
    # candidates from which (A^H r, |a|^2) run faster as the sample-domain scan
    ADJOINT_MIN_ATOMS = 128
    SCAN_MIN_ATOMS = 512
    SCAN_RANGE = 512  # apertures per range of the scan launch

    def correlate(self, params, residual, scan=True):
        """(a^H r, |a|^2) of every atom, area included: the sample-domain scan (one scan, on the
        matrix cores) where it serves the chirp and the candidates are many, else (and with
        scan=False) the correlation kernel, which correlate_gradient shares."""
        config, positions = self.sensor
        residual = residual.reshape(self.shape)
        if scan and len(params["centers"]) >= self.SCAN_MIN_ATOMS and surfel.kernel.sample_domain(config,
                                                                                                 self.fresnel_n):
            corr, norm = surfel.kernel.scan_batched(config, [(positions, residual)], params["centers"],
                                                    params["normals"], self.fresnel_n, self.SCAN_RANGE)  # fmt: skip
            root = params["areas"].sqrt()
            return corr[0] * root, norm[0] * params["areas"]
        return surfel.correlate(config, positions, params, residual, self.fresnel_n)

    def adjoint_correlation(self, params, residual):
        """A^H r of every atom (correlate's first output) for the right-hand sides: through the
        sample-domain VJP kernel when it serves the chirp and the atoms are many, else the
        correlation kernel. Both are FP32-accurate (within ~1e-6 of float64)."""
        config = self.scan.fmcw_config
        if len(params["centers"]) >= self.ADJOINT_MIN_ATOMS and surfel.kernel.sample_domain(config, self.fresnel_n):
            return surfel.adjoint_correlation(*self.sensor, params, residual.reshape(self.shape), self.fresnel_n)
        return self.correlate(params, residual)[0]

    def correlate_gradient(self, params, residual, include_normals=False):
        return surfel.correlate_gradient(*self.sensor, params, residual.reshape(self.shape), self.fresnel_n,
                                         include_normals)  # fmt: skip

    def correlate_orientations(self, params, residual, bank, table=None):
        return surfel.correlate_orientations(*self.sensor, params, bank, residual.reshape(self.shape), self.fresnel_n,
                                             table)  # fmt: skip

    def orient_select(self, params, residual, bank, table=None, real_part=False):
        return surfel.orient_select(*self.sensor, params, bank, residual.reshape(self.shape), self.fresnel_n, table,
                                    real_part)  # fmt: skip

    @property
    def lockstep_search(self):
        """Whether qn_search runs in lockstep rounds of matrix-core evaluations (chirps the
        sample-domain kernels serve), which fill the chip at any candidate count."""
        return surfel.kernel.sample_domain(self.sensor[0], self.fresnel_n)

    def qn_search(self, params, residual, spec, lockstep=None):
        """Quasi-Newton certificate ascent of every candidate on the device: lockstep rounds
        (lockstep_search by default) or one launch with a workgroup per candidate."""
        config, positions = self.sensor
        residual = residual.reshape(self.shape)
        if self.lockstep_search if lockstep is None else lockstep:
            operands = surfel.kernel.CertificateOperands(config, residual, self.fresnel_n)
            return search.qn_lockstep(positions, params["centers"], params["normals"], params["areas"], spec, operands)
        return search.qn_search(config, positions, residual, params["centers"], params["normals"], params["areas"],
                                spec, self.fresnel_n)  # fmt: skip

    @torch.no_grad()
    def range_gate(self, points, measurement, relative_power, margin_cells, min_fraction):
        """Points whose spherical range falls in a measured range interval at min_fraction of
        the apertures with signal. 32 evenly spaced apertures mark the bins whose power reaches
        relative_power of their own peak, dilated by margin_cells range resolutions
        (fmcw.py:312-335)."""
        config, positions = self.sensor
        probes = torch.linspace(0, len(measurement) - 1, min(32, len(measurement)), device=points.device).long()
        power = measurement[probes].abs().square()
        peak = power.max(-1).values
        lit = peak > peak.max() * 1e-6
        power, peak = power[lit], peak[lit]
        spacing = config.range_bin_spacing_m
        reach = math.ceil(margin_cells * config.range_resolution_m / spacing)
        marked = (power >= relative_power * peak[:, None]).float()[:, None]
        intervals = F.max_pool1d(marked, 2 * reach + 1, stride=1, padding=reach)[:, 0].int()
        return search.range_gate(points, positions[probes][lit], intervals, spacing, min_fraction)

    @torch.no_grad()
    def range_table(self, residual, oversample=8):
        """Range correlations of the residual at n_fft * oversample positions: the residual's
        sample-domain coefficients, zero-padded and transformed, and the norm of an atom at
        each position (the observed-bin mask circularly convolved with the kernel's power)
        (fmcw.py:337-355)."""
        config = self.scan.fmcw_config
        n, samples = config.n_fft, config.adc_samples
        size = n * oversample
        profiles = residual.reshape(self.shape)
        coefficients = torch.fft.ifft(profiles, n=n, dim=-1)[..., :samples] * (n / samples)
        correlation = torch.fft.fft(coefficients, n=size, dim=-1)
        response = torch.fft.fft(profiles.real.new_ones(samples), n=size) / samples
        observed = profiles.real.new_zeros(size)
        observed[: config.num_range_bins * oversample : oversample] = 1
        norm = torch.fft.ifft(torch.fft.fft(observed) * torch.fft.fft(response.abs().square())).real
        return RangeTable(torch.view_as_real(correlation).contiguous(), norm.contiguous(), oversample)


@dataclass(frozen=True)
class FMCWModel(_SplatProducts, _LeastSquaresStream, _SearchSupport):
    """Surfels measured by an FMCW scan (fmcw.py:129-355 of the reference).

    params = {centers [S, 3], normals [S, 3], areas [S]}; the coefficients are the
    reflectivities. prepare returns a copy that carries the range subspace.
    """

    scan: Scan
    fresnel_n: float = 0.0
    basis: RangeSubspace | None = None
    bands: tuple | None = None  # per aperture, on the host: lowest and highest peak bin
    block_bases: dict = field(default_factory=dict, repr=False)
    split_cache: SplitCache | None = field(default=None, repr=False, compare=False)

    @property
    def shape(self):
        return tuple(self.scan.data.shape)

    @property
    def size(self):
        return self.scan.data.numel()

    @property
    def apertures(self):
        return self.shape[0]

    @property
    def row_alignment(self):
        return self.shape[1]

    @property
    def sensor(self):
        return self.scan.fmcw_config, self.scan.scan_positions_m

    def sample_apertures(self, count, seed):
        """Sorted ids of a seeded random set of whole apertures, and the model on them."""
        ids = torch.randperm(self.shape[0], generator=torch.Generator().manual_seed(seed))[: max(1, count)]
        ids = ids.sort().values.to(self.scan.data.device)
        subset = replace(self.scan, data=self.scan.data[ids].contiguous(),
                         scan_positions_m=self.scan.scan_positions_m[ids].contiguous())  # fmt: skip
        return ids, replace(self, scan=subset, bands=None, block_bases={}, split_cache=None)

    def sample_rows(self, count, seed, device):
        """Every row of a seeded random set of count // K whole apertures."""
        bins = self.shape[1]
        chosen = torch.randperm(self.shape[0], generator=torch.Generator().manual_seed(seed))[: max(1, count // bins)]
        return (chosen.to(device)[:, None] * bins + torch.arange(bins, device=device)).flatten()
