"""Bindings of the solver kernels that replace host loops of small tensor operations."""

import torch

from .. import _lib


def backward_elimination(inverse, b, count, limited=None, limit=0):
    """Removed unknown indices [count] int64 of the on-device backward elimination;
    inverse [n, n] float64 real, b [n] = inverse q (replacement.py:104-139)."""
    n = len(inverse)
    removed = torch.zeros(count, dtype=torch.int64, device=inverse.device)
    if count == 0:
        return removed
    size = _lib.library().dsplat_elimination_workspace(n, count)
    if size == 0:
        raise ValueError(f"backward elimination kernel supports count <= n and count <= 4096; got {n}, {count}")
    workspace = torch.empty(size // 8 + 1, dtype=torch.float64, device=inverse.device)
    flags = None if limited is None else limited.to(torch.int32).contiguous()
    _lib.call(
        "dsplat_backward_elimination", _lib.ptr(inverse.contiguous()), _lib.ptr(b.contiguous()), n, count,
        _lib.ptr(flags), int(limit), _lib.ptr(removed), _lib.ptr(workspace), size, _lib.stream(inverse.device),
    )  # fmt: skip
    return removed


def _native(h):
    return h.is_cuda and h.dtype == torch.float64 and h.ndim == 2 and h.shape[0] == h.shape[1]


def cholesky_ex(h):
    """(L, info) like torch.linalg.cholesky_ex for a real float64 or float32 [n, n] device
    matrix: L lower with zeros above the diagonal, info a 0-d int32 tensor (first non-positive
    pivot, 1-based; 0 on success). Other dtypes and shapes go to torch.linalg.cholesky_ex."""
    single = h.is_cuda and h.dtype == torch.float32 and h.ndim == 2 and h.shape[0] == h.shape[1]
    if not (_native(h) or single):
        return torch.linalg.cholesky_ex(h)
    factor = h.contiguous().clone()
    info = torch.empty((), dtype=torch.int32, device=h.device)
    n = len(factor)
    entry = "dsplat_cholesky_f32" if single else "dsplat_cholesky"
    _lib.call(entry, _lib.ptr(factor), n, n, _lib.ptr(info), _lib.stream(h.device))
    return factor, info


def cholesky_solve(rhs, factor):
    """torch.cholesky_solve(rhs, factor) (lower factor) for real float64 [n, m] right-hand
    sides on the device: blocked forward and backward substitution, one launch per 64-row
    block and direction; other cases go to torch."""
    n = len(factor)
    if not (_native(factor) and rhs.dtype == torch.float64 and rhs.ndim == 2 and rhs.is_cuda):
        return torch.cholesky_solve(rhs, factor)
    # [m, n], a copy (the kernel works in place): column j of the solution is row j
    x = rhs.t().clone(memory_format=torch.contiguous_format)
    size = _lib.library().dsplat_cholesky_solve_workspace(n, rhs.shape[1])
    workspace = torch.empty(max(size // 8, 1), dtype=torch.float64, device=factor.device)
    _lib.call(
        "dsplat_cholesky_solve", _lib.ptr(factor.contiguous()), n, n, _lib.ptr(x), rhs.shape[1], n,
        _lib.ptr(workspace), size, _lib.stream(factor.device),
    )  # fmt: skip
    return x.t()


def power_iteration(h, start, steps):
    """The last product y = h x of `steps` power-iteration steps from `start` (x normalized by
    max(|y|, tiny) between steps) for a real float64 [n, n] device matrix, or None when the
    kernel does not apply (varpro.py:31-41)."""
    n = len(h)
    if not (_native(h) and 1 <= n <= 16384 and steps >= 1):
        return None
    h = h.contiguous()
    out = torch.empty(n, dtype=torch.float64, device=h.device)
    work = torch.empty(2 * n, dtype=torch.float64, device=h.device)
    _lib.call(
        "dsplat_power_iteration", _lib.ptr(h), n, n, int(steps), _lib.ptr(start.to(torch.float64).contiguous()),
        _lib.ptr(out), _lib.ptr(work), _lib.stream(h.device),
    )  # fmt: skip
    return out


def pcg_start(rhs, diagonal, tolerance):
    """(x, residual, direction, state) of conjugate gradients from zero on float32 device
    vectors: x = 0, residual = rhs, direction = rhs / diagonal; state [4] float32 = (rz,
    running flag, goal = tolerance |rhs|, 0) (varpro.py:206-214)."""
    rhs, diagonal = rhs.float().contiguous(), diagonal.float().contiguous()
    x, residual, direction = (torch.empty_like(rhs) for _ in range(3))
    state = rhs.new_empty(4)
    _lib.call(
        "dsplat_pcg_start", _lib.ptr(rhs), _lib.ptr(diagonal), len(rhs), float(tolerance), _lib.ptr(x),
        _lib.ptr(residual), _lib.ptr(direction), _lib.ptr(state), _lib.stream(rhs.device),
    )  # fmt: skip
    return x, residual, direction, state


def pcg_step(image, diagonal, x, residual, direction, state):
    """One conjugate-gradient iteration in place after image = H direction, the stopping
    tests on the device (varpro.py:215-228); a stopped iteration changes nothing."""
    _lib.call(
        "dsplat_pcg_step", _lib.ptr(image.float().contiguous()), _lib.ptr(diagonal), len(x), _lib.ptr(x),
        _lib.ptr(residual), _lib.ptr(direction), _lib.ptr(state), _lib.stream(x.device),
    )  # fmt: skip


def joint_layout(tail, complex_mode, centers=None, normals=None, areas=None):
    """dsplat_joint_layout of the joint system's unknowns: each trainable field as (first
    unknown, scale, unit vectors), None when it is not trainable; the coefficients from tail."""
    layout = _lib.JointLayout(centers=-1, normals=-1, areas=-1, tail=int(tail), complex_mode=int(complex_mode))
    for key, entry in (("centers", centers), ("normals", normals), ("areas", areas)):
        if entry is None:
            continue
        first, scale, unit = entry
        setattr(layout, key, int(first))
        setattr(layout, key[:-1] + "_scale", float(scale))
        if key != "areas":
            setattr(layout, key + "_unit", int(unit))
    return layout


def joint_unpack(vector, layout, centers, normals, coefficients, root, coefficient_scale):
    """(dcenters, dnormals, dweights complex) of the dimensionless unknowns vector
    (refinement.py:136-146 with the weight chain of surfel.py:53-61), one launch."""
    count = len(centers)
    dc = torch.empty_like(centers)
    dn = torch.empty_like(normals)
    dw = torch.empty((count, 2), dtype=torch.float32, device=centers.device)
    _lib.call(
        "dsplat_joint_unpack", count, _lib.ptr(vector), _lib.ctypes.byref(layout), _lib.ptr(centers),
        _lib.ptr(normals), _lib.ptr(torch.view_as_real(coefficients)), _lib.ptr(root), _lib.ptr(coefficient_scale),
        _lib.ptr(dc), _lib.ptr(dn), _lib.ptr(dw), _lib.stream(centers.device),
    )  # fmt: skip
    return dc, dn, torch.view_as_complex(dw)


def joint_pack(gcenters, gnormals, gweights, layout, centers, normals, coefficients, root, coefficient_scale, vector,
               damping, normalization):
    """The operator image adjoint / normalization + damping * vector from the VJP's gradients
    (refinement.py:148-158, 170-174), one launch."""
    out = torch.empty_like(vector)
    _lib.call(
        "dsplat_joint_pack", len(centers), _lib.ptr(gcenters.contiguous()), _lib.ptr(gnormals.contiguous()),
        _lib.ptr(torch.view_as_real(gweights.contiguous())), _lib.ctypes.byref(layout), _lib.ptr(centers),
        _lib.ptr(normals), _lib.ptr(torch.view_as_real(coefficients)), _lib.ptr(root), _lib.ptr(coefficient_scale),
        _lib.ptr(vector), _lib.ptr(damping), float(normalization), _lib.ptr(out), _lib.stream(centers.device),
    )  # fmt: skip
    return out

