"""The C ABI library: it loads, exports every declared symbol, and rejects bad input on the host.

No kernel is launched here: rejected configurations return before any GPU call.
"""

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "dsplat_hip.h"


def declared():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"\b(dsplat_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "dsplat_surfel_render" in names and "dsplat_surfel_vjp" in names


def test_library_exports_every_declared_symbol():
    from dsplat import _lib

    lib = _lib.library()
    missing = [name for name in declared() if not hasattr(lib, name)]
    assert not missing, missing
    # the Python binding declares a signature for every entry point of the header
    assert sorted(_lib.SIGNATURES) == declared()
    assert b"gfx950" in lib.dsplat_version()


def test_rejected_configurations_return_invalid_value_without_launching():
    from dsplat import FMCWConfig, _lib

    lib = _lib.library()
    bad = _lib.fmcw(FMCWConfig())
    bad.pad = 3  # pad factors outside {1, 2, 4, 8} are rejected
    code = lib.dsplat_surfel_render(ctypes.byref(bad), None, 1, None, None, None, 1, None, None)
    assert code == 1
    assert b"invalid" in lib.dsplat_error_string(code).lower()
    bad = _lib.fmcw(FMCWConfig())
    bad.num_bins = 100  # K must be n_fft / 2
    args = [None, 1, None, None, None, 1, None, None, None, None, None, 0, None]
    assert lib.dsplat_surfel_vjp(ctypes.byref(bad), *args) == 1
    # the top selection holds at most DSPLAT_TOP_MAX (4096) in LDS and takes at least one
    for top in (0, 4097):
        assert lib.dsplat_top_descending(None, 1, 10, top, None, None) == 1
    assert lib.dsplat_top_descending(None, 1, 10, 256, None, None) == 1  # null buffers
    assert lib.dsplat_top_descending(None, 0, 10, 256, None, None) == 0  # nothing to rank
    assert lib.dsplat_certificate_scores(None, 1, 10, None, None, 0, None, None) == 1
    # the lockstep ascent serves sample-domain chirps only, and checks its spec and workspace
    samples, bins = _lib.fmcw(FMCWConfig()), _lib.fmcw(FMCWConfig(splat_form="bins"))
    assert lib.dsplat_qn_lockstep_workspace(ctypes.byref(bins), 64, 16) == 0
    size = lib.dsplat_qn_lockstep_workspace(ctypes.byref(samples), 64, 16)
    assert size > 0
    spec = _lib.QnSpec()
    spec.width, spec.iterations, spec.backtracks = 6, 4, 2
    fake = ctypes.c_void_p(256)  # never dereferenced: the checks come first
    args = [fake, 64, fake, fake, fake, fake, 16, ctypes.byref(spec), fake, fake, fake, fake]
    assert lib.dsplat_qn_lockstep(ctypes.byref(samples), *args, size - 1, None) == 1  # short workspace
    assert lib.dsplat_qn_lockstep(ctypes.byref(bins), *args, size, None) == 1  # bin-domain chirp
    spec.width = 4
    assert lib.dsplat_qn_lockstep(ctypes.byref(samples), *args, size, None) == 1  # width 3 or 6 only
    spec.width = 6
    assert lib.dsplat_qn_lockstep(ctypes.byref(samples), *args[:6], 0, *args[7:], size, None) == 0  # no candidates


def test_vjp_workspace_per_splat_form():
    from dsplat import FMCWConfig, _lib

    lib = _lib.library()
    cfg = ctypes.byref(_lib.fmcw(FMCWConfig()))
    assert lib.dsplat_surfel_vjp_workspace(cfg, 0, 10) == 0
    # sample-domain form at 128 and 256 samples (pad >= 2): partials of 8 floats per (surfel,
    # 64-aperture range), one slot of 32 bins x 10 complex moments per aperture (2560 bytes)
    # and its band (int2)
    moments = 2048 * 256 * 8 * 4 + 16384 * 2560 + 16384 * 8
    assert lib.dsplat_surfel_vjp_workspace(cfg, 16384, 2048) == moments
    wide = ctypes.byref(_lib.fmcw(FMCWConfig(adc_samples=256)))
    assert lib.dsplat_surfel_vjp_workspace(wide, 16384, 2048) == moments
    # pad 1 keeps the FP16 hi + lo MFMA form: partials per 128-aperture range, the symmetric
    # split's operands (one K half: 2 KiB per aperture) and one scale per aperture
    pad1 = ctypes.byref(_lib.fmcw(FMCWConfig(pad_factor=1)))
    assert lib.dsplat_surfel_vjp_workspace(pad1, 16384, 2048) == 2048 * 128 * 8 * 4 + 16384 * 2048 + 16384 * 4
    # bin-domain form, chosen per call: one partial per surfel and 32-aperture block
    bins = ctypes.byref(_lib.fmcw(FMCWConfig(splat_form="bins")))
    assert lib.dsplat_surfel_vjp_workspace(bins, 16384, 2048) == 2048 * 512 * 8 * 4
    bad = _lib.fmcw(FMCWConfig())
    bad.splat_form = 2  # forms other than 0 and 1 are rejected
    assert lib.dsplat_surfel_vjp_workspace(ctypes.byref(bad), 16384, 2048) == 0
    odd = _lib.fmcw(FMCWConfig(adc_samples=96))  # no sample-domain kernel: bin-domain sizes
    assert lib.dsplat_surfel_vjp_workspace(ctypes.byref(odd), 1024, 200) == 200 * 32 * 8 * 4


def test_struct_layout_matches_the_header():
    from dsplat import _lib

    # six int32 (the chirp sizes, the splat form, a reserved word) then seven doubles,
    # naturally aligned: 24 + 56 bytes
    assert ctypes.sizeof(_lib.Fmcw) == 80
    assert _lib.Fmcw.splat_form.offset == 16
    assert _lib.Fmcw.k0_per_m.offset == 24


def test_kernels_refuse_host_tensors():
    import torch

    from dsplat import FMCWConfig
    from dsplat.kernels import surfel

    with pytest.raises(ValueError, match="device"):
        surfel._profiles(FMCWConfig(), torch.zeros(1, 3), torch.zeros(1, 3), torch.zeros(1, 3), torch.zeros(1, 2), 0.0)
