"""Dirichlet atoms on a spectral grid: the signal-level measurement model.

Public surface of the reference's dsplat/spectral.py:16-120 (SpectralGrid, SpectralModel).
An atom has a centre [D] in bin units and a complex coefficient; its spectrum is the
separable windowed Dirichlet kernel of every axis, observed at every bin of the grid.

The bin lattice of a grid is made once per device and kept resident. Correlations a^H r
run as the spectral VJP kernel (the coefficient gradient of Re<r, A b> is A^H r), so no
dictionary block is materialized for them; atoms that sit exactly on the grid bins use the
FFT shortcut instead (the correlation is a circular cross-correlation then).
"""

from dataclasses import dataclass
from math import prod

import torch

from .kernels import spectral as kernel

_LATTICES: dict = {}


def _lattice(shape, device):
    """Every bin's coordinates [size, D] (row-major), float32, cached per (shape, device)."""
    key = (tuple(shape), str(torch.device(device)))
    if key not in _LATTICES:
        axes = torch.meshgrid(*(torch.arange(n, device=device, dtype=torch.float32) for n in shape), indexing="ij")
        _LATTICES[key] = torch.stack(axes, -1).reshape(-1, len(shape))
    return _LATTICES[key]


@dataclass(frozen=True)
class SpectralGrid:
    shape: tuple[int, ...]  # FFT length per axis (one to three axes)
    samples: tuple[int, ...] | None = None  # observed samples per axis; None means the FFT length
    window: str = "rect"  # rect, hann, hamming or kaiser
    beta: float = 14.0  # Kaiser shape parameter
    radius: float = 0.0  # support in bins; 0 keeps the exact, untruncated response

    def __post_init__(self):
        if self.samples is None:
            object.__setattr__(self, "samples", self.shape)

    @property
    def size(self):
        return prod(self.shape)

    def bins(self, device="cuda"):
        """Every bin's coordinates [size, D], row-major."""
        return _lattice(self.shape, device).clone()


# Dictionary entries per streamed block of the norm pass of truncated atoms.
_NORM_BLOCK = 2**24


@dataclass(frozen=True)
class SpectralModel:
    """Atoms with centres [S, D] on one grid: render, jvp, vjp, design, jacobian, streamed
    blocks, correlations and row sampling (spectral.py:37-120 of the reference). A grid has no
    apertures, so the solver's aperture-subset options never apply to it."""

    grid: SpectralGrid
    row_alignment = 1
    apertures = 0

    @property
    def shape(self):
        return self.grid.shape

    @property
    def size(self):
        return self.grid.size

    def prepare(self, specs):
        return self

    def _bins(self, params):
        return _lattice(self.grid.shape, params["centers"].device)

    def design(self, params, rows):
        return kernel.design(self.grid, params["centers"], self._bins(params)[rows])

    def jacobian(self, params, rows, atoms, keys):
        _, derivative = kernel.design(self.grid, params["centers"][atoms], self._bins(params)[rows], derivatives=True)
        return {"centers": derivative}

    def render(self, params, coefficients):
        return kernel.render(self.grid, params["centers"], coefficients, self._bins(params)).reshape(self.shape)

    def jvp(self, params, coefficients, tangent, delta_coefficients):
        bins = self._bins(params)
        flat = kernel.jvp(self.grid, params["centers"], coefficients, tangent["centers"], delta_coefficients, bins)
        return flat.reshape(self.shape)

    def vjp(self, params, coefficients, upstream):
        by_center, by_coefficient = kernel.vjp(self.grid, params["centers"], coefficients, upstream.flatten(),
                                               self._bins(params))  # fmt: skip
        return {"centers": by_center}, by_coefficient

    def _norms(self, centers):
        """|a_s|^2 per atom. Untruncated atoms observed on the whole grid all have the same
        energy (Parseval over the periodic kernel), taken from one atom; truncated ones are
        summed over streamed dictionary blocks."""
        bins = _lattice(self.grid.shape, centers.device)
        if self.grid.radius <= 0:
            one = kernel.design(self.grid, centers[:1], bins)
            return one.abs().square().sum().expand(len(centers)).clone()
        norms = centers.new_zeros(len(centers))
        step = max(1, _NORM_BLOCK // max(1, len(centers)))
        for first in range(0, self.size, step):
            norms += kernel.design(self.grid, centers, bins[first : first + step]).abs().square().sum(0)
        return norms

    def correlate(self, params, residual, scan=True):
        """a^H r and |a|^2 for every atom (spectral.py:90-106 of the reference); scan is
        accepted for the FMCW model's signature and has no effect here."""
        centers = params["centers"]
        bins = self._bins(params)
        flat = residual.reshape(-1)
        if centers.shape == bins.shape and torch.equal(centers, bins):
            # atoms on every bin: corr[c] = sum_k conj(atom(k - c)) r(k), a circular
            # cross-correlation with the atom at the origin
            origin = kernel.design(self.grid, torch.zeros_like(centers[:1]), bins).reshape(self.shape)
            spectrum = torch.fft.fftn(flat.reshape(self.shape)) * torch.fft.fftn(origin).conj()
            energy = origin.abs().square().sum()
            return torch.fft.ifftn(spectrum).reshape(-1), energy.expand(len(centers))
        # the coefficient gradient of Re<r, A b> is A^H r, whatever b is
        _, corr = kernel.vjp(self.grid, centers, torch.ones(len(centers), dtype=flat.dtype, device=flat.device),
                             flat, bins)  # fmt: skip
        return corr, self._norms(centers)

    def adjoint_correlation(self, params, residual):
        """A^H r of every atom (correlate's first output)."""
        return self.correlate(params, residual)[0]

    def blocks(self, params, target, real, chunk):
        """Least-squares blocks (design rows, observations) of `chunk` bins; the real form
        stacks the real and imaginary rows of each bin."""
        flat = target.reshape(-1)
        for first in range(0, self.size, chunk):
            rows = slice(first, first + chunk)
            a, y = self.design(params, rows), flat[rows]
            if real:
                a = torch.view_as_real(a).transpose(1, 2).reshape(-1, a.shape[1])
                y = torch.view_as_real(y).reshape(-1)
            yield a, y

    def sample_rows(self, count, seed, device):
        """Evenly spaced bins (the grid is deterministic; seed is unused)."""
        return torch.linspace(0, self.size - 1, min(count, self.size), device=device).long()
