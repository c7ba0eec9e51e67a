"""Synthetic ground truth: planar arrays and float64 FMCW scans of known surfels
(groundtruth.py:16-90 of the reference)."""

import numpy as np
import torch


def _turn_z_onto(axis):
    """The rotation [3, 3] (float64) that takes +z onto the unit vector axis by the shortest
    arc: angle acos(z . axis) about z x axis (a half turn about x when axis is -z)."""
    z = np.array([0.0, 0.0, 1.0])
    pivot = np.cross(z, axis)
    sine, cosine = np.linalg.norm(pivot), float(z @ axis)
    if sine < 1e-10:
        return np.eye(3) if cosine > 0 else np.diag([1.0, -1.0, -1.0])
    k = pivot / sine
    cross_k = np.array([[0.0, -k[2], k[1]], [k[2], 0.0, -k[0]], [-k[1], k[0], 0.0]])
    return np.eye(3) + sine * cross_k + (1 - cosine) * (cross_k @ cross_k)


def planar_array(side, pitch, axis=(0.0, 0.0, 1.0), standoff=0.35, *, device="cpu"):
    """side x side aperture positions [side^2, 3] in metres, pitch apart, row-major
    (groundtruth.py:16-37 of the reference).

    The grid is centred on the origin in the plane z = 0 and faces along axis toward the
    point (0, 0, standoff): the plane turns about that point by the shortest rotation taking
    +z onto axis (no turn for the default +z).
    """
    direction = np.asarray(axis, dtype=np.float64)
    turn = _turn_z_onto(direction / np.linalg.norm(direction))
    offsets = (np.arange(side) - (side - 1) / 2) * pitch
    u, v = np.meshgrid(offsets, offsets, indexing="ij")
    flat = np.column_stack([u.ravel(), v.ravel(), np.zeros(side * side)])
    pivot = np.array([0.0, 0.0, standoff])
    return torch.tensor((flat - pivot) @ turn.T + pivot, dtype=torch.float32, device=device)


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
