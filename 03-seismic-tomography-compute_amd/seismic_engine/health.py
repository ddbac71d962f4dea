"""Field health functions and status types of the reference's module
(src/cpp/pybind/bindings.cpp:275-301 and its enums), computed on the GPU.

Arrays may be numpy (copied to the device and back; 1-D float64 results
returned like the reference's forcecast arrays) or torch tensors (worked on
in place on their device)."""
from __future__ import annotations

import ctypes as C
import enum
from dataclasses import dataclass

import numpy as np

from ._lib import call, ptr, stream_handle


class HealthStatus(enum.IntEnum):      # src/cpp/core/fault_tolerance.h:13-20
    HEALTHY = 0
    NAN_DETECTED = 1
    INF_DETECTED = 2
    CFL_VIOLATION = 3
    DIVERGENCE = 4
    COMM_FAILURE = 5


class StepResult(enum.IntEnum):        # src/cpp/solver/wave_equation.h:24-29
    OK = 0
    NAN_DETECTED_AND_REPAIRED = 1
    CFL_VIOLATION = 2
    FATAL_DIVERGENCE = 3


@dataclass
class PropagationResult:               # wave_equation.h:31-37
    status: StepResult = StepResult.OK
    total_nan_encountered: int = 0
    total_cells_repaired: int = 0
    timesteps_completed: int = 0
    completed_successfully: bool = False


@dataclass
class FaultStats:                      # fault_tolerance.h:61-67
    total_nan_count: int = 0
    total_inf_count: int = 0
    cells_repaired: int = 0
    timesteps_recovered: int = 0
    global_abort_count: int = 0


def _torch():
    import torch
    return torch


def _device_f64(data):
    """(device tensor, numpy array to write back or None)."""
    torch = _torch()
    if torch.is_tensor(data):
        if data.dtype != torch.float64 or not data.is_cuda or not data.is_contiguous():
            raise TypeError("tensor inputs must be contiguous float64 CUDA tensors (they are modified in place)")
        return data.reshape(-1), None
    a = np.ascontiguousarray(np.asarray(data, dtype=np.float64)).reshape(-1)
    return torch.as_tensor(a).to("cuda"), a


def check_field_health(data) -> int:
    """Number of NaN/Inf values (FaultTolerance::count_nan_inf)."""
    d, _ = _device_f64(data)
    n = C.c_int64()
    call("stp_count_nonfinite_f64", ptr(d), C.c_int64(d.numel()), C.byref(n), stream_handle())
    return n.value


def clamp_field_safe(data, lo: float = -1e10, hi: float = 1e10):
    """Clamp to [lo, hi], non-finite values to 0 (FaultTolerance::clamp_field)."""
    d, host = _device_f64(data)
    call("stp_clamp_field_f64", ptr(d), C.c_int64(d.numel()), C.c_double(lo), C.c_double(hi), stream_handle())
    return _write_back(data, d, host)


def _write_back(data, d, host):
    """Copy the device result into the host array; like the reference's
    forcecast binding, a float64 C-contiguous input is updated in place and
    returned itself."""
    if host is None:
        return data
    host[:] = d.cpu().numpy()
    if isinstance(data, np.ndarray) and np.shares_memory(host, data):
        return data
    return host


def repair_field_nan(data, nx: int, ny: int, nz: int, halo_width: int = 1, replace_value: float = 0.0):
    """Replace non-finite interior values by the mean of their finite interior
    6-neighbours (else replace_value), then any remaining non-finite value by
    replace_value (FaultTolerance::repair_field_nan, fault_tolerance.cpp:98-157).
    nx, ny, nz are the with-halo extents of the flat x-major array."""
    d, host = _device_f64(data)
    call("stp_repair_field_nan_f64", ptr(d), C.c_int64(d.numel()), C.c_int64(nx), C.c_int64(ny), C.c_int64(nz),
         int(halo_width), C.c_double(replace_value), stream_handle())
    return _write_back(data, d, host)
