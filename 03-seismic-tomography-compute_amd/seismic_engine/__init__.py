"""seismic_engine — MI355X (gfx950) drop-in for the reference's pybind11 module
of the same name (src/cpp/pybind/bindings.cpp:13).

The names the reference's Python code imports from it (gradient.py:5,
model.py:6, inversion.py:12-15) are here: CSRMatrix, SparseOps,
compute_fd_stencil, check_field_health, clamp_field_safe, repair_field_nan,
StepResult, HealthStatus, plus the RTM and result types.  All compute runs in
libseismic_hip.so through its C ABI; torch provides device memory and
streams.  Importing this module does not touch the GPU; the first compute
call loads the library and fails loudly if it is missing.
"""
from ._lib import SeismicError, lib  # noqa: F401
from . import synthetic  # noqa: F401
from .fd import FDPropagator  # noqa: F401
from .health import (FaultStats, HealthStatus, PropagationResult, StepResult, check_field_health,  # noqa: F401
                     clamp_field_safe, repair_field_nan)
from .rtm import (RTMConfig, RTMEngine, RTMImageCondition, SliceAxis, SliceSpec,  # noqa: F401
                  compute_adjoint_gradient)
from .sparse import CSRMatrix, DeviceCSR, SparseOps, compute_fd_stencil  # noqa: F401

__all__ = ["SeismicError", "lib", "synthetic", "FDPropagator", "FaultStats", "HealthStatus", "PropagationResult",
           "StepResult", "check_field_health", "clamp_field_safe", "repair_field_nan", "RTMConfig", "RTMEngine",
           "RTMImageCondition", "SliceAxis", "SliceSpec", "compute_adjoint_gradient", "CSRMatrix", "DeviceCSR",
           "SparseOps", "compute_fd_stencil"]
