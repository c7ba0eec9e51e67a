"""FMCW sensor: chirp configuration, scans, and surfels measured as range profiles.

Same public surface as dsplat/fmcw.py of the reference. A linear chirp of
bandwidth B and carrier fc is sampled by adc_samples ADC samples and zero-padded to
n_fft; a scan holds the positive range bins of every aperture,
data[A, K] = FFT(ADC, n_fft)[:K] / adc_samples. A surfel at range r appears at bin
r / spacing, spacing = c / (2 B pad_factor) (fmcw.py:1-13).
"""

from dataclasses import dataclass

import torch


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
