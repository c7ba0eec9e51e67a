"""Synthetic ground truth: planar arrays and float64 FMCW scans of known surfels
(groundtruth.py:16-90 of the reference)."""

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

    Runs as one float64 matrix-core kernel (dsplat_synthesize) on the device of the
    surfels, which must be a GPU; geometry is taken as float32, as the reference's callers
    pass it. aperture_block and surfel_block are the reference's memory tiling and are
    accepted for compatibility: the kernel tiles on its own.
    """
    from .fmcw import Scan
    from .kernels import surfel as kernels

    device = surfel_centers.device
    spectrum = kernels.synthesize(
        config,
        aperture_positions.to(device, torch.float32).contiguous(),
        surfel_centers.to(torch.float32).contiguous(),
        surfel_normals.to(torch.float32).contiguous(),
        surfel_amplitudes.to(device),
        fresnel_n,
    )
    return Scan(spectrum.to(torch.complex64), aperture_positions.to(device), config)
