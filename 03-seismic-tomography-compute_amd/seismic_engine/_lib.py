"""ctypes binding of libseismic_hip.so (the C ABI declared in include/seismic_hip.h).

The library must exist: there is no fallback path.  Every call checks the
returned status and raises SeismicError with the library's message.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent.parent
LIB_PATH = Path(os.environ.get("STP_LIB", PKG_DIR / "lib" / "libseismic_hip.so"))


class SeismicError(RuntimeError):
    pass


class FdConfig(C.Structure):
    _fields_ = [
        ("nx", C.c_int64), ("ny", C.c_int64), ("nz", C.c_int64),
        ("halo_width", C.c_int32), ("order", C.c_int32),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("dt", C.c_double),
        ("sponge_width", C.c_int32), ("sponge_damping", C.c_double),
        ("amplitude_limit", C.c_double), ("cfl_safety_factor", C.c_double),
        ("nshots", C.c_int32),
    ]


class SolveInfo(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32), ("rel_residual", C.c_double),
                ("true_rel_residual", C.c_double), ("replacements", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_vp = C.c_void_p
_i32 = C.c_int32
_i64 = C.c_int64
_f64 = C.c_double

# name -> argtypes (all return int status unless listed in _RESTYPES)
SIGNATURES: dict[str, list] = {
    "stp_version": [],
    "stp_device_alloc": [C.POINTER(_vp), _i64],
    "stp_device_free": [_vp],
    "stp_memcpy_h2d": [_vp, _vp, _i64, _vp],
    "stp_memcpy_d2h": [_vp, _vp, _i64, _vp],
    "stp_memcpy_d2d": [_vp, _vp, _i64, _vp],
    "stp_memset_d": [_vp, C.c_int, _i64, _vp],
    "stp_stream_sync": [_vp],
    "stp_device_count": [C.POINTER(C.c_int)],
    "stp_fd_plan_create": [C.POINTER(FdConfig), C.POINTER(_vp)],
    "stp_fd_plan_destroy": [_vp],
    "stp_fd_set_receivers": [_vp, _vp, _i32],
    "stp_fd_set_sources": [_vp, _vp, _i32],
    "stp_fd_set_sponge_mode": [_vp, _i32],
    "stp_fd_coefficients": [_vp, _vp, _vp, C.POINTER(_f64), _vp],
    "stp_fd_verify_cfl": [_vp, _f64, C.POINTER(_i32)],
    "stp_fd_propagate": [_vp, _vp, _vp, _vp, _i32, _i32, _vp, _i32, _i32, _vp, _i32, _vp, _vp],
    "stp_fd_sponge_factors": [_vp, _vp, _vp],
    "stp_stream_3r1w_f32": [_i64, _vp, _vp, _vp, _vp, _i32, _vp],
    "stp_field_energy_f32": [_vp, _i64, C.POINTER(_f64), _vp],
    "stp_field_interior_f64": [_vp, _i64, _i64, _i64, _i32, _vp, _vp],
    "stp_count_nonfinite_f64": [_vp, _i64, C.POINTER(_i64), _vp],
    "stp_clamp_field_f64": [_vp, _i64, _f64, _f64, _vp],
    "stp_repair_field_nan_f64": [_vp, _i64, _i64, _i64, _i64, _i32, _f64, _vp],
    "stp_csr_create": [_i64, _i64, _i64, _vp, _vp, _vp, C.POINTER(_vp), _vp],
    "stp_csr_from_rays": [_i64, _vp, _f64, _vp, C.POINTER(_vp), _vp],
    "stp_csr_destroy": [_vp],
    "stp_csr_shape": [_vp, C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_i64)],
    "stp_csr_export": [_vp, _i32, _vp, _vp, _vp, _vp],
    "stp_csr_spmv": [_vp, _i32, _vp, _vp, _vp],
    "stp_csr_spmv_f64": [_vp, _i32, _vp, _vp, _vp],
    "stp_csr_set_grid": [_vp, _i64, _i64, _i64],
    "stp_csr_prepare_layout": [_vp, _vp],
    "stp_csr_staging": [_vp, C.POINTER(_i32), C.POINTER(_i64)],
    "stp_csr_set_values_f64": [_vp, _vp, _vp],
    "stp_csr_gather_probe": [_vp, _i32, _i32, _vp, _vp, _vp],
    "stp_cg": [_vp, _vp, _i32, _f64, _i32, _vp, _vp, C.POINTER(SolveInfo), _vp],
    "stp_cgls": [_vp, _vp, _f64, _f64, _i32, _i32, _vp, _vp, _vp, C.POINTER(SolveInfo), _vp],
    "stp_cgls_noalloc": [_vp, _vp, _f64, _f64, _i32, _i32, _vp, _vp, _vp, C.POINTER(SolveInfo), _vp],
    "stp_csr_assemble": [_i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, C.POINTER(_i64), _vp],
    "stp_csr64_fd7_triplets": [_i64, _i64, _i64, _vp, _f64, _f64, _f64, _f64, _vp, _vp, _vp, _vp],
    "stp_csr64_scale": [_i64, _vp, _f64, _vp],
    "stp_csr64_add_diagonal": [_i64, _vp, _vp, _vp, _f64, _vp],
    "stp_dot_f64": [_i64, _vp, _vp, C.POINTER(_f64), _vp],
    "stp_csr64_to_f32": [_i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "stp_fd_adjoint": [_vp, _vp, _vp, _i32, _i32, _vp, _vp, _i32, _i32, _i32, _vp, _vp, _vp, C.POINTER(_f64), _vp],
    "stp_rtm_image_condition_f64": [_i64, _i64, _i64, _i32, _i64, _i64, _vp, _vp, _vp, _vp, _vp],
    "stp_rtm_finalize_f64": [_i64, _vp, _vp, _vp, _vp],
    "stp_accumulate_f32_f64": [_i64, _vp, _vp, _vp],
    "stp_smooth_gradient_f64": [_i64, _i64, _i64, _f64, _vp, _vp, _vp, _vp],
    "stp_gaussian_filter_3d_f64": [_i64, _i64, _i64, _f64, _vp, _vp, _vp],
    "stp_adjoint_gradient_f64": [_i64, _vp, _vp, _vp, _f64, _vp, _vp],
    "stp_lpc_create": [C.c_uint64, C.POINTER(_vp)],
    "stp_lpc_destroy": [_vp],
    "stp_lpc_num_params": [C.POINTER(_i64)],
    "stp_lpc_get_params": [_vp, _vp, _i64],
    "stp_lpc_set_params": [_vp, _vp, _i64],
    "stp_lpc_features": [_vp, _vp, _vp, _f64, _vp, _vp, _vp],
    "stp_lpc_prepare": [_vp, _vp, _vp, _f64, _vp],
    "stp_lpc_prepare_train": [_vp, _vp, _vp, _f64, _vp],
    "stp_lpc_coefficients": [_vp, _vp, _vp],
    "stp_lpc_apply": [_vp, _vp, _vp, _vp],
    "stp_lpc_backward_accumulate": [_vp, _vp, _vp, _vp],
    "stp_lpc_backward": [_vp, _f64, _vp, _vp],
    "stp_lpc_adam_step": [_vp, _f64, _vp],
    "stp_conv3d_bf16": [_i64, _i64, _i64, _i32, _i32, _vp, _vp, _i32, _vp, _vp],
    "stp_conv3d_bf16_bwd": [_i64, _i64, _i64, _i32, _i32, _vp, _vp, _vp, _vp, _vp, _vp],
    "stp_tt_forward": [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "stp_tt_gradient": [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "stp_gradient_tikhonov_f64": [_i64, _i64, _i64, _f64, _vp, _vp, _vp],
    "stp_gradient_normalize_f64": [_i64, _vp, C.POINTER(_f64), _vp],
    "stp_depth_preconditioner_f64": [_i64, _i64, _i64, _i64, _f64, _vp, _vp],
    "stp_gradient_precondition_f64": [_i64, _vp, _vp, _vp],
    "stp_terrain_bathymetry_synthetic_f64": [_i64, _i64, _f64, _f64, _f64, _f64, _f64, _vp, _vp],
    "stp_terrain_velocity_mask_f64": [_vp, _vp, _vp, _f64, _f64, _vp, _vp],
    "stp_terrain_boundary_mask": [_vp, _vp, _vp, _vp, _vp],
    "stp_terrain_free_surface_f64": [_i64, _i64, _i64, _i64, _vp, _vp],
    "stp_terrain_seafloor_f64": [_vp, _vp, _vp, _vp, _i64, _vp],
    "stp_terrain_absorbing_f64": [_i64, _i64, _i64, _i64, _i32, _f64, _vp, _vp],
    "stp_rtm_extract_slice_f64": [_i32, _i64, _i64, _i64, _i64, _i32, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp,
                                  _vp],
}
_RESTYPES = {"stp_last_error": C.c_char_p}

_lib = None


def lib() -> C.CDLL:
    """Load the library once; raise if it has not been built."""
    global _lib
    if _lib is None:
        # torch ships its own HIP runtime with the same soname as ROCm's; load
        # it first so that this library binds to the same runtime (one set of
        # streams and device allocations in the process)
        import torch  # noqa: F401
        if not LIB_PATH.exists():
            raise SeismicError(
                f"{LIB_PATH} is missing: build it with `python __graft_entry__.py build` "
                "(there is no CPU fallback)")
        h = C.CDLL(str(LIB_PATH))
        h.stp_last_error.restype = C.c_char_p
        h.stp_last_error.argtypes = []
        for name, args in SIGNATURES.items():
            fn = getattr(h, name)
            fn.argtypes = args
            fn.restype = C.c_int
        _lib = h
    return _lib


def check(status: int, what: str = "") -> None:
    if status != 0:
        msg = lib().stp_last_error().decode(errors="replace")
        raise SeismicError(f"{what or 'seismic_hip'} failed ({status}): {msg}")


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args), name)


def ptr(t) -> int:
    """Device/host address of a torch tensor (0 for None)."""
    return 0 if t is None else int(t.data_ptr())


def stream_handle(device=None) -> int:
    import torch
    return int(torch.cuda.current_stream(device).cuda_stream)
