"""FMCW sensor: chirp configuration, scans, and surfels measured as range profiles.

Same public surface as dsplat/fmcw.py of the reference. A linear chirp of
bandwidth B and carrier fc is sampled by adc_samples ADC samples and zero-padded to
n_fft; a scan holds the positive range bins of every aperture,
data[A, K] = FFT(ADC, n_fft)[:K] / adc_samples. A surfel at range r appears at bin
r / spacing, spacing = c / (2 B pad_factor) (fmcw.py:1-13).
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
    """Chirp constants (fmcw.py:29-65)."""

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
    """Range profiles data [A, K] measured at scan positions [A, 3], metres (fmcw.py:68-74)."""

    data: torch.Tensor
    scan_positions_m: torch.Tensor
    fmcw_config: FMCWConfig


# --- Row basis of the range response (fmcw.py:77-118) -------------------------------------

PEAKS_PER_BIN = 16  # samples of the peak position per range bin for the response SVD


@dataclass(frozen=True)
class RowBasis:
    """Orthonormal row projection for observations stored as consecutive aperture blocks."""

    adjoint: torch.Tensor  # complex [rank, bins]; U^H
    interleaved: torch.Tensor  # real [2 rank, 2 bins] acting on (re, im) row pairs

    def reduce(self, values):
        """Complex rows [blocks * bins, ...] to [blocks * rank, ...]."""
        tail = values.shape[1:]
        rows = values.reshape(-1, self.adjoint.shape[1], tail.numel())
        return (self.adjoint @ rows).reshape(-1, *tail)

    def reduce_major(self, rows):
        """Real rows ordered bin by bin, (2 k + part) * blocks + block, in one product."""
        tail = rows.shape[1:]
        return (self.interleaved @ rows.reshape(self.interleaved.shape[1], -1)).reshape(-1, *tail)


def response_basis(config, low, high, device):
    """Row basis of every peak position in [low, high] bins, or None if it spans all bins.

    The float64 responses come from the range_response kernel; the SVD runs on the host in
    float64 and keeps singular values above eps32 / 128 of the largest, so the normal
    equations are preserved to float32 precision (fmcw.py:101-118).
    """
    count = max(2 * config.adc_samples, int(PEAKS_PER_BIN * (high - low)) + 1)
    response = range_response(config, low, high, count, device).cpu()
    vectors, values, _ = torch.linalg.svd(response, full_matrices=False)
    rank = int((values > torch.finfo(torch.float32).eps / 128 * values[0]).sum())
    if rank >= config.num_range_bins:
        return None
    adjoint = vectors[:, :rank].mH
    interleaved = torch.zeros((2 * rank, 2 * config.num_range_bins), dtype=torch.float64)
    interleaved[0::2, 0::2], interleaved[0::2, 1::2] = adjoint.real, -adjoint.imag
    interleaved[1::2, 0::2], interleaved[1::2, 1::2] = adjoint.imag, adjoint.real
    return RowBasis(adjoint.to(device, torch.complex64), interleaved.to(device, torch.float32))


class RangeTable(NamedTuple):
    """FFT-interpolated range correlations of a residual; they only rank proposals."""

    correlation: torch.Tensor  # [apertures, n_fft * oversample, 2]
    norm: torch.Tensor  # [n_fft * oversample]; keeps the observed-bin mask
    oversample: int


# --- Measurement model (fmcw.py:129-355) ----------------------------------------------------


@dataclass(frozen=True)
class FMCWModel:
    """Surfels measured by an FMCW scan: the measurement model the solver works with.

    params = {centers [S, 3], normals [S, 3], areas [S]}; the coefficients are
    reflectivities. render, jvp, vjp, design and jacobian on chosen rows, streamed
    least-squares blocks, correlations, seeded row and aperture sampling, the range
    gate and the FFT range table. prepare returns a copy with the row basis.
    """

    scan: Scan
    fresnel_n: float = 0.0
    basis: RowBasis | None = None
    bands: tuple | None = None  # per aperture, on the host: lowest and highest peak bin
    block_bases: dict = field(default_factory=dict, repr=False)

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
        generator = torch.Generator().manual_seed(seed)
        ids = torch.randperm(self.shape[0], generator=generator)[: max(1, count)]
        ids = ids.sort().values.to(self.scan.data.device)
        scan = replace(
            self.scan,
            data=self.scan.data[ids].contiguous(),
            scan_positions_m=self.scan.scan_positions_m[ids].contiguous(),
        )
        return ids, replace(self, scan=scan, bands=None, block_bases={})

    def sample_rows(self, count, seed, device):
        generator = torch.Generator().manual_seed(seed)
        bins = self.shape[1]
        apertures = torch.randperm(self.shape[0], generator=generator)[: max(1, count // bins)]
        return (apertures.to(device)[:, None] * bins + torch.arange(bins, device=device)).flatten()

    def prepare(self, specs):
        """This model with the row basis of every center inside the box of the centers'
        bounds; the band is the nearest and farthest box point of any aperture, one bin
        of margin (fmcw.py:196-218)."""
        lower, upper = specs["centers"].lower, specs["centers"].upper
        config, positions = self.sensor
        low, high, bands = 0.0, float(config.n_fft), None
        if lower is not None and upper is not None:
            lo, hi = positions.new_tensor(lower).expand(3), positions.new_tensor(upper).expand(3)
            nearest = torch.minimum(torch.maximum(positions, lo), hi)
            farthest = torch.where((positions - lo).abs() > (positions - hi).abs(), lo, hi)
            spacing = config.range_bin_spacing_m
            nearest = ((positions - nearest).norm(dim=1) / spacing - 1).clamp(min=0)
            farthest = (positions - farthest).norm(dim=1) / spacing + 1
            near, far = float(nearest.min()), float(farthest.max())
            if far - near < config.n_fft:
                low, high = near, far
                bands = (nearest.cpu().numpy(), farthest.cpu().numpy())
        basis = response_basis(config, low, high, positions.device)
        return replace(self, basis=basis, bands=bands, block_bases={})

    def block_basis(self, rows):
        """Row basis for a block of whole apertures, at most as large as the shared one
        (fmcw.py:220-241)."""
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
            basis = response_basis(config, float(key[0]), float(key[1]), positions.device)
            smaller = basis is not None and len(basis.adjoint) < len(self.basis.adjoint)
            self.block_bases[key] = basis if smaller else self.basis
        return self.block_bases[key]

    def blocks(self, params, target, real, chunk):
        """Least-squares blocks of whole apertures, projected onto the row basis; real
        blocks are ordered bin by bin so one product projects them (fmcw.py:243-261)."""
        config, positions = self.sensor
        bins = self.shape[1]
        flat = target.flatten()
        for start in range(0, self.size, chunk):
            rows = slice(start, min(start + chunk, self.size))
            basis = self.block_basis(rows)
            if real:
                block = positions[rows.start // bins : rows.stop // bins]
                a = surfel.design(config, block, params, self.fresnel_n, real=True)
                y = torch.view_as_real(flat[rows]).reshape(len(block), 2 * bins).T.reshape(-1)
                yield (a, y) if basis is None else (basis.reduce_major(a), basis.reduce_major(y))
            else:
                a, y = self.design(params, rows), flat[rows]
                yield (a, y) if basis is None else (basis.reduce(a), basis.reduce(y))

    def projected_blocks(self, params, chunk, column_max=False):
        """Real row-projected design blocks of whole apertures on the shared row basis, atoms
        generated on chip and never written (the blocks of `blocks` without y), or None
        when there is no basis to project on. With column_max, (block, maxima) pairs: the
        float bits of every column's largest |entry| (int32 [S]) where the sample-domain
        projection forms them on the way, else None."""
        if self.basis is None:
            return None
        config, positions = self.sensor
        bins = self.shape[1]
        apertures = max(1, chunk // bins)
        basis = self.basis.interleaved
        sampled = surfel.kernel.sample_basis(config, basis, self.fresnel_n)  # once per pass
        count = len(params["centers"])

        def block(a0):
            maxima = None
            if column_max and sampled is not None:
                maxima = torch.zeros(count, dtype=torch.int32, device=basis.device)
            a = surfel.kernel.project_design(
                config, positions[a0 : a0 + apertures], params["centers"], params["normals"], params["areas"], basis,
                self.fresnel_n, sampled, maxima,
            )  # fmt: skip
            return (a, maxima) if column_max else a

        return (block(a0) for a0 in range(0, len(positions), apertures))

    def _apertures(self, rows):
        apertures, inverse = torch.unique(rows // self.shape[1], return_inverse=True)
        return self.scan.scan_positions_m[apertures], inverse, rows % self.shape[1]

    def design(self, params, rows):
        """Complex dictionary rows [rows, S]; a slice must cover whole apertures."""
        config, positions = self.sensor
        bins = self.shape[1]
        if isinstance(rows, slice):
            block = positions[rows.start // bins : -(-rows.stop // bins)]
            return surfel.design(config, block, params, self.fresnel_n)
        block, inverse, ids = self._apertures(rows)
        a = surfel.design(config, block, params, self.fresnel_n)
        return a.reshape(len(block), bins, -1)[inverse, ids]

    def jacobian(self, params, rows, atoms, keys):
        block, inverse, ids = self._apertures(rows)
        selected = {key: value[atoms] for key, value in params.items()}
        derivatives = surfel.jacobian(self.scan.fmcw_config, block, selected, self.fresnel_n)
        return {key: derivatives[key][inverse, ids] for key in keys}

    def render(self, params, coefficients):
        return surfel.render(*self.sensor, params, coefficients, self.fresnel_n)

    def jvp(self, params, coefficients, tangent, delta_coefficients):
        return surfel.jvp(*self.sensor, params, coefficients, tangent, delta_coefficients, self.fresnel_n)

    def vjp(self, params, coefficients, upstream):
        return surfel.vjp(*self.sensor, params, coefficients, upstream, self.fresnel_n)

    def correlate(self, params, residual):
        return surfel.correlate(*self.sensor, params, residual.reshape(self.shape), self.fresnel_n)

    def correlate_gradient(self, params, residual, include_normals=False):
        return surfel.correlate_gradient(
            *self.sensor, params, residual.reshape(self.shape), self.fresnel_n, include_normals
        )

    def correlate_orientations(self, params, residual, bank, table=None):
        return surfel.correlate_orientations(
            *self.sensor, params, bank, residual.reshape(self.shape), self.fresnel_n, table
        )

    def qn_search(self, params, residual, spec):
        """Quasi-Newton certificate ascent of every candidate in one launch."""
        config, positions = self.sensor
        return search.qn_search(
            config, positions, residual.reshape(self.shape), params["centers"], params["normals"], params["areas"], spec,
            self.fresnel_n,
        )  # fmt: skip

    def orient_select(self, params, residual, bank, table=None, real_part=False):
        return surfel.orient_select(
            *self.sensor, params, bank, residual.reshape(self.shape), self.fresnel_n, table, real_part
        )

    @torch.no_grad()
    def range_gate(self, points, measurement, relative_power, margin_cells, min_fraction):
        """Points whose spherical range falls in a measured range interval at enough
        apertures: 32 evenly spaced apertures threshold their profile relative to its
        peak, dilate by margin_cells resolutions, and gate (fmcw.py:312-335)."""
        config, positions = self.sensor
        count = min(32, len(measurement))
        ids = torch.linspace(0, len(measurement) - 1, count, device=points.device).long()
        power = measurement[ids].abs().square()
        peak = power.max(-1).values
        informative = peak > peak.max() * 1e-6
        power, peak = power[informative], peak[informative]
        spacing = config.range_bin_spacing_m
        radius = math.ceil(margin_cells * config.range_resolution_m / spacing)
        above = (power >= relative_power * peak[:, None]).float()[:, None]
        intervals = F.max_pool1d(above, 2 * radius + 1, stride=1, padding=radius)[:, 0].int()
        return search.range_gate(points, positions[ids][informative], intervals, spacing, min_fraction)

    @torch.no_grad()
    def range_table(self, residual, oversample=8):
        """Oversampled FFT range correlations of the residual for proposal scores
        (fmcw.py:337-355)."""
        config = self.scan.fmcw_config
        n, samples = config.n_fft, config.adc_samples
        size = n * oversample
        residual = residual.reshape(self.shape)
        coefficients = torch.fft.ifft(residual, n=n, dim=-1)[..., :samples] * (n / samples)
        correlation = torch.fft.fft(coefficients, n=size, dim=-1)
        kernel = torch.fft.fft(residual.real.new_ones(samples), n=size) / samples
        mask = residual.real.new_zeros(size)
        mask[: config.num_range_bins * oversample : oversample] = 1
        norm = torch.fft.ifft(torch.fft.fft(mask) * torch.fft.fft(kernel.abs().square())).real
        return RangeTable(torch.view_as_real(correlation).contiguous(), norm.contiguous(), oversample)
