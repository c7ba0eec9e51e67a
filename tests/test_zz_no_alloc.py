"""dsplat_surfel_vjp_noalloc (the VJP on the library's own scratch array) against
dsplat_surfel_vjp with a caller-allocated workspace, through the C ABI.

Two calls in a row, the second smaller, in both splat forms, each compared bit for bit with
the workspace entry point (same kernels, no atomics, deterministic); a size beyond the
scratch is refused on the host before any launch.
"""

import ctypes
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def say(*args):
    print(*args, file=sys.stderr, flush=True)


def scene(apertures, surfels, seed):
    g = torch.Generator().manual_seed(seed)
    side = int(np.ceil(np.sqrt(apertures)))
    t = (torch.arange(side) - (side - 1) / 2) * 0.0025
    x, y = torch.meshgrid(t, t, indexing="ij")
    pos = torch.stack([x.flatten(), y.flatten(), torch.zeros(side * side)], -1)[:apertures]
    centers = torch.tensor([0.0, 0.0, 0.35]) + 0.03 * (2 * torch.rand(surfels, 3, generator=g) - 1)
    normals = torch.nn.functional.normalize(torch.randn(surfels, 3, generator=g) + torch.tensor([0, 0, -2.0]), dim=1)
    weights = torch.randn(surfels, dtype=torch.complex64, generator=g) * 1e-3
    up = torch.randn(apertures, 128, dtype=torch.complex64, generator=g)
    d = torch.device("cuda", 0)
    return [v.to(d).contiguous() for v in (pos, centers, normals, torch.view_as_real(weights), torch.view_as_real(up))]


def call_both(cfg, pos, c, n, w, up):
    from dsplat import _lib

    lib = _lib.library()
    A, S = len(pos), len(c)
    st = _lib.stream(pos.device)
    outs = []
    for noalloc in (False, True):
        gc, gn, gw = torch.full((S, 3), 7.0, device=pos.device), torch.full((S, 3), 7.0, device=pos.device), \
            torch.full((S, 2), 7.0, device=pos.device)  # fmt: skip
        if noalloc:
            code = lib.dsplat_surfel_vjp_noalloc(ctypes.byref(cfg), _lib.ptr(pos), A, _lib.ptr(c), _lib.ptr(n),
                                                 _lib.ptr(w), S, _lib.ptr(up), _lib.ptr(gc), _lib.ptr(gn),
                                                 _lib.ptr(gw), st)  # fmt: skip
        else:
            size = lib.dsplat_surfel_vjp_workspace(ctypes.byref(cfg), A, S)
            ws = torch.empty(size // 4 + 1, device=pos.device)
            code = lib.dsplat_surfel_vjp(ctypes.byref(cfg), _lib.ptr(pos), A, _lib.ptr(c), _lib.ptr(n), _lib.ptr(w),
                                         S, _lib.ptr(up), _lib.ptr(gc), _lib.ptr(gn), _lib.ptr(gw), _lib.ptr(ws),
                                         size, st)  # fmt: skip
        assert code == 0, lib.dsplat_error_string(code)
        torch.cuda.synchronize()
        outs.append([x.cpu().numpy() for x in (gc, gn, gw)])
    return outs


@pytest.mark.parametrize("form", ["samples", "bins"])
def test_noalloc_vjp_matches_the_workspace_entry_point_on_two_calls(form):
    from dsplat import FMCWConfig, _lib

    cfg = _lib.fmcw(FMCWConfig(splat_form=form), 0.0)
    for apertures, surfels, seed in ((4096, 1024, 1), (300, 77, 2)):
        say(f"no-alloc VJP, {form} form: {apertures} apertures x {surfels} surfels (seed {seed})")
        ref, got = call_both(cfg, *scene(apertures, surfels, seed))
        for a, b in zip(ref, got):
            assert np.isfinite(a).all()
            assert np.array_equal(a, b)


def test_noalloc_vjp_refuses_a_workspace_beyond_its_scratch():
    from dsplat import FMCWConfig, _lib

    lib = _lib.library()
    cfg = _lib.fmcw(FMCWConfig(), 0.0)
    limit = lib.dsplat_noalloc_scratch_bytes()
    assert limit == 64 << 20
    # 16384 x 2048 at 128 samples fits; 65536 apertures of 4096 surfels do not
    assert lib.dsplat_surfel_vjp_workspace(ctypes.byref(cfg), 16384, 2048) <= limit
    assert lib.dsplat_surfel_vjp_workspace(ctypes.byref(cfg), 65536, 4096) > limit
    d = torch.device("cuda", 0)
    dummy = torch.zeros(16, device=d)
    p = _lib.ptr(dummy)
    code = lib.dsplat_surfel_vjp_noalloc(ctypes.byref(cfg), p, 65536, p, p, p, 4096, p, p, p, p, _lib.stream(d))
    assert code == 2  # hipErrorOutOfMemory, returned before any launch
