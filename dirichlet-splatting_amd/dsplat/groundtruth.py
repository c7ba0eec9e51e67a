"""Synthetic ground truth: planar arrays and float64 FMCW scans of known surfels
(groundtruth.py:16-90 of the reference)."""

import math

import numpy as np
import torch


def planar_array(side, pitch, axis=(0.0, 0.0, 1.0), standoff=0.35, *, device="cpu"):
    """side x side aperture positions [side^2, 3] in metres, pitch apart, row-major.

    The array faces along axis toward (0, 0, standoff); the default is the plane z = 0
    facing +z. The rotation maps +z onto axis by the Rodrigues formula.
    """
    axis = np.asarray(axis, dtype=np.float64)
    axis = axis / np.linalg.norm(axis)
    source = np.array([0.0, 0.0, 1.0])
    cross = np.cross(source, axis)
    dot = float(source @ axis)
    if np.linalg.norm(cross) < 1e-10:
        rotation = np.eye(3) if dot > 0 else np.diag([1.0, -1.0, -1.0])
    else:
        x, y, z = cross
        skew = np.array([[0, -z, y], [z, 0, -x], [-y, x, 0]])
        rotation = np.eye(3) + skew + skew @ skew / (1 + dot)
    t = (np.arange(side) - (side - 1) / 2) * pitch
    gx, gy = np.meshgrid(t, t, indexing="ij")
    base = np.stack([gx.ravel(), gy.ravel(), np.zeros(side * side)], -1)
    center = np.array([0.0, 0.0, standoff])
    return torch.tensor(center + (base - center) @ rotation.T, dtype=torch.float32, device=device)


@torch.no_grad()
def synthesize(
    aperture_positions, surfel_centers, surfel_normals, surfel_amplitudes, config, fresnel_n=0.0,
    *, aperture_block=256, surfel_block=512,
):  # fmt: skip
    """A Scan of surfels with complex amplitudes (reflectivity times sqrt(area)), made by
    the FMCW signal chain in float64: each aperture-surfel echo is a beat tone of frequency
    2 slope r / c0 sampled by the ADC, the tones are summed, zero-padded to n_fft, FFT'd
    and cut to the positive bins, divided by adc_samples (groundtruth.py:58-90). It never
    evaluates the Dirichlet kernel, so a fit to its scans does not reuse the forward model.
    """
    from .fmcw import Scan

    device = surfel_centers.device
    c = config
    t = torch.arange(c.adc_samples, device=device, dtype=torch.float64) / c.sample_rate_hz
    centers, normals = surfel_centers.double(), surfel_normals.double()
    amplitudes = surfel_amplitudes.to(torch.complex128)
    rows = []
    for a0 in range(0, len(aperture_positions), aperture_block):
        block = aperture_positions[a0 : a0 + aperture_block].to(device, torch.float64)
        adc = torch.zeros((len(block), 1, c.adc_samples), dtype=torch.complex128, device=device)
        for s0 in range(0, len(centers), surfel_block):
            sl = slice(s0, s0 + surfel_block)
            offset = block[:, None, :] - centers[None, sl]
            r = offset.norm(dim=-1).clamp(min=1e-9)
            cosine = ((offset / r[..., None]) * normals[None, sl]).sum(-1).clamp(min=0)
            gain = cosine.square() * (c.lambda_m / (8 * math.pi)) / r
            if fresnel_n > 1:
                root = (fresnel_n * fresnel_n - 1 + cosine.square()).sqrt()
                gain = gain * (root - cosine) / (root + cosine)
            delay = 2 * r / c.c0_m_s
            carrier = 2 * math.pi * (c.fc_hz * delay + c.slope_hz_s * delay * (c.adc_start_time_s - 0.5 * delay))
            echo = torch.polar(gain, carrier) * amplitudes[sl]
            beat = 2 * c.slope_hz_s * r / c.c0_m_s
            tones = torch.polar(torch.ones_like(r)[..., None], 2 * math.pi * beat[..., None] * t)
            adc += torch.bmm(echo[:, None, :], tones)
        spectrum = torch.fft.fft(adc[:, 0], n=c.n_fft)[:, : c.num_range_bins]
        rows.append((spectrum / c.adc_samples).to(torch.complex64))
    return Scan(torch.cat(rows), aperture_positions.to(device), config)
