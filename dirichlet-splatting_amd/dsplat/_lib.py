"""ctypes binding of libdsplat_hip.so, the C ABI declared in include/dsplat_hip.h.

torch is imported first on purpose: its wheel ships the HIP runtime
(libamdhip64.so.7), and the library must bind to that same runtime so that torch's
device pointers and streams are valid for our launches. There is no fallback: if the
library is missing every kernel call raises.
"""

import ctypes
import os
from functools import cache
from pathlib import Path

import torch

LIB_PATH = Path(os.environ.get("DSPLAT_LIB", Path(__file__).parent / "_lib" / "libdsplat_hip.so"))

c_int, c_size, c_double, c_float = ctypes.c_int32, ctypes.c_size_t, ctypes.c_double, ctypes.c_float
c_ptr = ctypes.c_void_p


class Fmcw(ctypes.Structure):
    """dsplat_fmcw of include/dsplat_hip.h."""

    _fields_ = [
        ("num_bins", c_int),
        ("n_fft", c_int),
        ("samples", c_int),
        ("pad", c_int),
        ("splat_form", c_int),
        ("reserved", c_int),
        ("k0_per_m", c_double),
        ("fc_hz", c_double),
        ("slope_hz_s", c_double),
        ("t_start_s", c_double),
        ("lambda_m", c_double),
        ("c0_m_s", c_double),
        ("fresnel_n", c_double),
    ]


# name -> (restype, argtypes); one line per entry point of include/dsplat_hip.h
class JointLayout(ctypes.Structure):
    """dsplat_joint_layout of dsplat_hip.h."""

    _fields_ = [
        ("centers", ctypes.c_int32),
        ("normals", ctypes.c_int32),
        ("areas", ctypes.c_int32),
        ("tail", ctypes.c_int32),
        ("complex_mode", ctypes.c_int32),
        ("centers_unit", ctypes.c_int32),
        ("normals_unit", ctypes.c_int32),
        ("center_scale", ctypes.c_float),
        ("normal_scale", ctypes.c_float),
        ("area_scale", ctypes.c_float),
    ]


class QnSpec(ctypes.Structure):
    """dsplat_qn_spec of dsplat_hip.h."""

    _fields_ = [
        ("center_scale", ctypes.c_float),
        ("normal_scale", ctypes.c_float),
        ("center_mask", ctypes.c_float * 3),
        ("normal_mask", ctypes.c_float * 3),
        ("direction_mask", ctypes.c_float * 6),
        ("step_scale", ctypes.c_float * 6),
        ("lower", ctypes.c_float * 3),
        ("upper", ctypes.c_float * 3),
        ("width", ctypes.c_int32),
        ("normalize", ctypes.c_int32),
        ("real_part", ctypes.c_int32),
        ("iterations", ctypes.c_int32),
        ("backtracks", ctypes.c_int32),
        ("radius", ctypes.c_float),
        ("power", ctypes.c_double),
    ]


SIGNATURES = {
    "dsplat_version": (ctypes.c_char_p, []),
    "dsplat_error_string": (ctypes.c_char_p, [c_int]),
    "dsplat_surfel_render": (c_int, [c_ptr, c_ptr, c_int, c_ptr, c_ptr, c_ptr, c_int, c_ptr, c_ptr]),
    "dsplat_surfel_jvp": (
        c_int, [c_ptr, c_ptr, c_int, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_int, c_ptr, c_ptr]
    ),
    "dsplat_surfel_vjp_workspace": (c_size, [c_ptr, c_int, c_int]),
    "dsplat_cholesky": (c_int, [c_ptr, c_int, c_int, c_ptr, c_ptr]),
    "dsplat_cholesky_f32": (c_int, [c_ptr, c_int, c_int, c_ptr, c_ptr]),
    "dsplat_cholesky_solve_workspace": (c_size, [c_int, c_int]),
    "dsplat_cholesky_solve": (c_int, [c_ptr, c_int, c_int, c_ptr, c_int, c_int, c_ptr, c_size, c_ptr]),
    "dsplat_power_iteration": (c_int, [c_ptr, c_int, c_int, c_int, c_ptr, c_ptr, c_ptr, c_ptr]),
    "dsplat_joint_unpack": (c_int, [c_int, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr,
                                    c_ptr]),
    "dsplat_joint_pack": (c_int, [c_int, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr,
                                  c_ptr, ctypes.c_float, c_ptr, c_ptr]),
    "dsplat_surfel_vjp": (
        c_int,
        [c_ptr, c_ptr, c_int, c_ptr, c_ptr, c_ptr, c_int, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_size, c_ptr],
    ),
    "dsplat_noalloc_scratch_bytes": (c_size, []),
    "dsplat_surfel_vjp_noalloc": (
        c_int, [c_ptr, c_ptr, c_int, c_ptr, c_ptr, c_ptr, c_int, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr]
    ),
    "dsplat_surfel_design": (c_int, [c_ptr, c_ptr, c_int, c_ptr, c_ptr, c_ptr, c_int, c_int, c_ptr, c_ptr]),
    "dsplat_surfel_jacobian": (c_int, [c_ptr, c_ptr, c_int, c_ptr, c_ptr, c_int, c_ptr, c_ptr, c_ptr]),
    "dsplat_surfel_correlate_workspace": (c_size, [c_int, c_int, c_int]),
    "dsplat_surfel_correlate": (
        c_int, [c_ptr, c_ptr, c_int, c_ptr, c_ptr, c_int, c_ptr, c_ptr, c_ptr, c_size, c_ptr]
    ),
    "dsplat_surfel_correlate_gradient": (
        c_int, [c_ptr, c_ptr, c_int, c_ptr, c_ptr, c_int, c_ptr, c_int, c_ptr, c_ptr, c_size, c_ptr]
    ),
    "dsplat_surfel_correlate_orientations": (
        c_int,
        [c_ptr, c_ptr, c_int, c_ptr, c_ptr, c_int, c_ptr, c_int, c_ptr, c_ptr, c_ptr, c_int, c_ptr, c_ptr, c_size,
         c_ptr],
    ),
    "dsplat_spectral_forward": (
        c_int, [c_int, c_ptr, c_ptr, c_int, c_float, c_ptr, c_int, c_ptr, c_ptr, c_int, c_ptr, c_int, c_ptr, c_ptr]
    ),
    "dsplat_spectral_jvp": (
        c_int,
        [c_int, c_ptr, c_ptr, c_int, c_float, c_ptr, c_int, c_ptr, c_ptr, c_ptr, c_ptr, c_int, c_ptr, c_int, c_ptr,
         c_ptr],
    ),
    "dsplat_spectral_design": (
        c_int, [c_int, c_ptr, c_ptr, c_int, c_float, c_ptr, c_int, c_ptr, c_int, c_ptr, c_int, c_ptr, c_ptr, c_ptr]
    ),
    "dsplat_spectral_backward_workspace": (c_size, [c_int, c_int, c_int]),
    "dsplat_spectral_backward": (
        c_int,
        [c_int, c_ptr, c_ptr, c_int, c_float, c_ptr, c_int, c_ptr, c_ptr, c_int, c_ptr, c_int, c_ptr, c_ptr, c_ptr,
         c_size, c_ptr],
    ),
    "dsplat_search_range_gate": (
        c_int, [c_ptr, c_int, c_ptr, c_int, c_ptr, c_int, c_float, c_int, c_ptr, c_ptr]
    ),
    "dsplat_search_peaks_workspace": (c_size, [c_int, c_int, c_int]),
    "dsplat_search_peaks": (c_int, [c_ptr, c_ptr, c_int, c_int, c_int, c_float, c_ptr, c_ptr, c_size, c_ptr]),
    "dsplat_search_direction": (c_int, [c_ptr, c_ptr, c_ptr, c_ptr, c_int, c_int, c_float, c_ptr, c_ptr]),
    "dsplat_search_trial": (c_int, [c_ptr, c_ptr, c_ptr, c_ptr, c_int, c_int, c_float, c_ptr, c_ptr, c_ptr]),
    "dsplat_energy_workspace": (c_size, [c_size]),
    "dsplat_energy": (c_int, [c_ptr, c_ptr, c_size, c_ptr, c_ptr, c_size, c_ptr]),
    "dsplat_range_response": (c_int, [c_ptr, c_double, c_double, c_int, c_int, c_int, c_int, c_ptr]),
    "dsplat_synthesize": (c_int, [c_ptr, c_ptr, c_int, c_ptr, c_ptr, c_ptr, c_int, c_ptr, c_ptr]),
    "dsplat_surfel_orient_select": (
        c_int,
        [c_ptr, c_ptr, c_int, c_ptr, c_ptr, c_ptr, c_int, c_ptr, c_int, c_ptr, c_ptr, c_ptr, c_int, c_int, c_ptr, c_ptr,
         c_ptr],
    ),
    "dsplat_project_design": (c_int, [c_ptr, c_ptr, c_int, c_ptr, c_ptr, c_ptr, c_int, c_ptr, c_int, c_ptr, c_ptr]),
    "dsplat_project_form": (c_int, [c_ptr, c_int]),
    "dsplat_sample_basis": (c_int, [c_ptr, c_ptr, c_int, c_ptr, c_ptr]),
    "dsplat_project_sampled": (
        c_int, [c_ptr, c_ptr, c_int, c_ptr, c_ptr, c_ptr, c_int, c_ptr, c_int, c_ptr, c_ptr, c_ptr]
    ),
    "dsplat_project_split_rows": (c_int, [c_int]),
    "dsplat_project_sampled_split": (
        c_int,
        [c_ptr, c_ptr, c_int, c_ptr, c_ptr, c_ptr, c_int, c_ptr, c_int, c_int, c_ptr, c_ptr, c_ptr, ctypes.c_int64,
         c_ptr],
    ),
    "dsplat_syrk_split_workspace": (c_size, [c_int, c_int, c_int]),
    "dsplat_syrk_split": (
        c_int, [c_int, c_ptr, c_ptr, ctypes.c_int64, c_int, c_int, c_int, c_ptr, c_ptr, c_ptr, c_size, c_ptr]
    ),
    "dsplat_qn_workspace": (c_size, [c_ptr, c_int]),
    "dsplat_qn_search": (
        c_int, [c_ptr, c_ptr, c_int, c_ptr, c_ptr, c_ptr, c_ptr, c_int, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_size, c_ptr]
    ),
    "dsplat_calibrate_valu": (c_int, [c_int, c_int, c_int, c_ptr, c_ptr]),
    "dsplat_qn_lockstep_workspace": (c_size, [c_ptr, c_int, c_int]),
    "dsplat_qn_lockstep": (
        c_int,
        [c_ptr, c_ptr, c_int, c_ptr, c_ptr, c_ptr, c_ptr, c_int, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_size, c_ptr],
    ),
    "dsplat_elimination_workspace": (c_size, [c_int, c_int]),
    "dsplat_backward_elimination": (c_int, [c_ptr, c_ptr, c_int, c_int, c_ptr, c_int, c_ptr, c_ptr, c_size, c_ptr]),
    "dsplat_syrk_workspace": (c_size, [c_int, c_int, c_int]),
    "dsplat_syrk": (c_int, [c_ptr, c_int, c_int, c_int, c_ptr, c_ptr, c_size, c_ptr]),
    "dsplat_syrk_scaled": (c_int, [c_ptr, c_int, c_int, c_int, c_ptr, c_ptr, c_ptr, c_size, c_ptr]),
    "dsplat_gemm_tn_workspace": (c_size, [c_int, c_int, c_int, c_int]),
    "dsplat_gemm_tn": (c_int, [c_ptr, c_int, c_int, c_int, c_int, c_ptr, c_ptr, c_size, c_ptr]),
    "dsplat_gemm_tn_scaled": (c_int, [c_ptr, c_int, c_int, c_int, c_int, c_ptr, c_ptr, c_ptr, c_size, c_ptr]),
    "dsplat_scan_workspace": (c_size, [c_ptr, c_int, c_ptr, c_int]),
    "dsplat_scan_batched": (
        c_int, [c_ptr, c_int, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_int, c_int, c_ptr, c_ptr, c_size, c_ptr]
    ),
    "dsplat_certificate_scores": (c_int, [c_ptr, c_int, c_int, c_ptr, c_ptr, c_int, c_ptr, c_ptr]),
    "dsplat_top_descending": (c_int, [c_ptr, c_int, c_int, c_int, c_ptr, c_ptr]),
    "dsplat_certificate_operands_bytes": (c_size, [c_ptr, c_int]),
    "dsplat_certificate_operands": (c_int, [c_ptr, c_ptr, c_int, c_ptr, c_size, c_ptr]),
    "dsplat_certificate_gradient_workspace": (c_size, [c_ptr, c_int, c_int]),
    "dsplat_certificate_gradient": (
        c_int, [c_ptr, c_ptr, c_int, c_ptr, c_ptr, c_int, c_ptr, c_ptr, c_ptr, c_size, c_ptr]
    ),
    "dsplat_pcg_start": (c_int, [c_ptr, c_ptr, c_int, c_float, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr]),
    "dsplat_pcg_step": (c_int, [c_ptr, c_ptr, c_int, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr]),
    "dsplat_gram_workspace": (c_size, [c_int, c_int]),
    "dsplat_gram": (
        c_int,
        [c_ptr, c_ptr, c_int, c_ptr, c_ptr, c_ptr, c_int, c_int, c_int, c_float, c_ptr, c_ptr, c_ptr, c_size, c_ptr],
    ),
}


@cache
def library():
    """The loaded library with every declared signature; raises if it is not built."""
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is not built: run `python dirichlet-splatting_amd/csrc/build.py` "
            "(or __graft_entry__.build()); dsplat has no CPU fallback."
        )
    lib = ctypes.CDLL(str(LIB_PATH))
    for name, (restype, argtypes) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype, fn.argtypes = restype, argtypes
    return lib


def call(name, *args):
    """Call an entry point and raise RuntimeError on a non-zero hipError_t."""
    lib = library()
    code = getattr(lib, name)(*args)
    if code:
        message = lib.dsplat_error_string(code).decode()
        raise RuntimeError(f"{name} failed with hipError {code}: {message}")


def ptr(tensor):
    """Device address of a contiguous tensor (None -> NULL)."""
    if tensor is None:
        return None
    if not tensor.is_cuda:
        raise ValueError("dsplat kernels take device tensors")
    if not tensor.is_contiguous():
        raise ValueError("dsplat kernels take contiguous tensors")
    return ctypes.c_void_p(tensor.data_ptr())


def stream(device):
    """The current HIP stream of a device as a raw hipStream_t."""
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


SPLAT_FORMS = {"samples": 0, "bins": 1}


def fmcw(config, fresnel_n=0.0):
    """The ABI struct of an FMCWConfig (kernels/surfel.py:21-32 sensor_args, plus fresnel_n)."""
    pad = int(config.pad_factor)
    if config.n_fft != config.adc_samples * pad:
        raise ValueError("n_fft must equal adc_samples * pad_factor")
    return Fmcw(
        num_bins=int(config.num_range_bins),
        n_fft=int(config.n_fft),
        samples=int(config.adc_samples),
        pad=pad,
        splat_form=SPLAT_FORMS[getattr(config, "splat_form", "samples")],
        k0_per_m=1.0 / config.range_bin_spacing_m,
        fc_hz=float(config.fc_hz),
        slope_hz_s=float(config.slope_hz_s),
        t_start_s=float(config.adc_start_time_s),
        lambda_m=float(config.lambda_m),
        c0_m_s=float(config.c0_m_s),
        fresnel_n=float(fresnel_n),
    )
