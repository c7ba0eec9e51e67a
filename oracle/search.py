"""Spatial-search primitives in plain numpy (oracle).

range_gate follows `dsplat/kernels/search.slang:7-24` (float32 distance, round half to
even); peaks is the brute-force greedy suppression of `tests/test_search.py:11-19`;
direction and trial follow `search.slang:68-109` in float64.
"""

import numpy as np


def range_gate(points, positions, intervals, spacing, required):
    points = np.asarray(points, np.float32)
    positions = np.asarray(positions, np.float32)
    d = points[:, None, :] - positions[None, :, :]
    x = np.sqrt((d * d).sum(-1, dtype=np.float32)).astype(np.float32) / np.float32(spacing)
    k = np.rint(x).astype(np.int64)  # numpy rounds half to even
    bins = intervals.shape[1]
    ok = (k >= 0) & (k < bins)
    hit = np.zeros_like(ok)
    rows = np.broadcast_to(np.arange(len(positions))[None], k.shape)
    hit[ok] = intervals[rows[ok], k[ok]] != 0
    return hit.sum(1) >= required


def peaks(scores, centers, count, separation):
    """Greedy best-first indices at least separation apart; ties to the lower index."""
    scores = np.array(scores, np.float64)
    centers = np.asarray(centers, np.float64)
    chosen = []
    for _ in range(min(count, len(scores))):
        finite = np.isfinite(scores)
        if not finite.any():
            break
        index = int(np.argmax(np.where(finite, scores, -np.inf)))
        if not np.isfinite(scores[index]):
            break
        chosen.append(index)
        d2 = ((centers - centers[index]) ** 2).sum(1)
        scores[d2 < separation * separation] = -np.inf
        scores[index] = -np.inf
    return chosen


def direction(inverse, gradient, normals, mask, radius):
    inverse, gradient = np.asarray(inverse, np.float64), np.asarray(gradient, np.float64)
    d = -np.einsum("cjk,ck->cj", inverse, gradient)
    descent = (d * gradient).sum(1)
    d = np.where((descent >= 0)[:, None], -gradient, d) * np.asarray(mask, np.float64)[None]
    if d.shape[1] == 6:
        n = np.asarray(normals, np.float64)
        d[:, 3:] -= (d[:, 3:] * n).sum(1, keepdims=True) * n
    length = np.sqrt((d * d).sum(1))
    return d * np.minimum(1.0, radius / np.maximum(length, 1e-30))[:, None]


def trial(centers, normals, direction, controls, fraction):
    c, n, dr, ctl = (np.asarray(x, np.float64) for x in (centers, normals, direction, controls))
    width = dr.shape[1]
    moved = c + fraction * ctl[0, :3] * dr[:, :3]
    out_c = np.where(ctl[1, :3] > 0, np.clip(moved, ctl[2, :3], ctl[3, :3]), c)
    if width == 6:
        m = n + fraction * ctl[0, 3:] * dr[:, 3:]
        out_n = m / np.sqrt((m * m).sum(1, keepdims=True))
    else:
        out_n = n
    return out_c, out_n
