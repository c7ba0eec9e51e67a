"""Synthetic ground truth: planar arrays (groundtruth.py:16-37 of the reference)."""

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
