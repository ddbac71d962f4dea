"""RTM imaging types of the reference's module (bindings.cpp:303-371) and an
RTMEngine with the reference's methods (src/cpp/rtm/rtm_engine.h:31-96):
FP64 image and illumination kept on the GPU.  Slice capture for live viewing
is not part of this build (specs are stored, no slice is produced)."""
from __future__ import annotations

import ctypes as C
import enum
from dataclasses import dataclass, field

import numpy as np

from ._lib import call, ptr, stream_handle


class SliceAxis(enum.IntEnum):
    X = 0
    Y = 1
    Z = 2
    VOLUME = 3


@dataclass
class SliceSpec:
    axis: SliceAxis = SliceAxis.Z
    index: int = -1
    stride: int = 1


@dataclass
class RTMConfig:                       # src/cpp/solver/wave_equation.h:39-48
    enable_rtm: bool = False
    enable_snapshots: bool = False
    snapshot_interval: int = 50
    enable_mmap_writer: bool = False
    mmap_filename: str = "rtm_snapshots.dat"
    mmap_max_bytes: int = 2 * 1024 ** 3
    ring_buffer_capacity: int = 1024
    slice_specs: list = field(default_factory=list)


@dataclass
class RTMImageCondition:
    image: np.ndarray
    illumination: np.ndarray
    reflectivity: np.ndarray


def _torch():
    import torch
    return torch


class RTMEngine:
    def __init__(self, nx: int, ny: int, nz: int, dx: float = 1.0, dy: float = 1.0, dz: float = 1.0,
                 halo_width: int = 1):
        torch = _torch()
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.spacing = (dx, dy, dz)
        self.halo_width = int(halo_width)
        n = self.total_cells()
        self.image = torch.zeros(n, dtype=torch.float64, device="cuda")
        self.illumination = torch.zeros_like(self.image)
        self.reflectivity = torch.zeros_like(self.image)
        self.slice_specs: list = []

    def total_cells(self) -> int:
        return self.nx * self.ny * self.nz

    def set_slice_specs(self, specs) -> None:
        self.slice_specs = list(specs)

    def apply_image_condition(self, forward_wavefield, adjoint_wavefield, ny: int, nz: int, step_idx: int = 0):
        """image += f a, illumination += f f over the interior (non-finite -> 0);
        inputs are with-halo arrays of y / z strides ny, nz."""
        torch = _torch()

        def dev(a):
            if torch.is_tensor(a):
                return a.to(device="cuda", dtype=torch.float64).contiguous().reshape(-1)
            return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64).reshape(-1)).to("cuda")
        f, a = dev(forward_wavefield), dev(adjoint_wavefield)
        call("stp_rtm_image_condition_f64", C.c_int64(self.nx), C.c_int64(self.ny), C.c_int64(self.nz),
             self.halo_width, C.c_int64(ny), C.c_int64(nz), ptr(f), ptr(a), ptr(self.image),
             ptr(self.illumination), stream_handle())

    def accumulate(self, image_f32, illumination_f32) -> None:
        """Add FP32 interior-indexed contributions (from a reverse-time pass)."""
        for src, dst in ((image_f32, self.image), (illumination_f32, self.illumination)):
            if src is not None:
                s = src.contiguous().reshape(-1)
                call("stp_accumulate_f32_f64", C.c_int64(s.numel()), ptr(s), ptr(dst), stream_handle())

    def finalize_image(self) -> None:
        call("stp_rtm_finalize_f64", C.c_int64(self.total_cells()), ptr(self.image), ptr(self.illumination),
             ptr(self.reflectivity), stream_handle())

    def reset_image(self) -> None:
        self.image.zero_()
        self.illumination.zero_()
        self.reflectivity.zero_()

    def result(self) -> RTMImageCondition:
        return RTMImageCondition(self.image.cpu().numpy(), self.illumination.cpu().numpy(),
                                 self.reflectivity.cpu().numpy())


def compute_adjoint_gradient(forward_wavefield, adjoint_wavefield, velocity, dt: float):
    """g = -fwd * adj * (2 v dt^2) elementwise, FP64 on the GPU
    (GradientComputer.compute_adjoint_gradient, gradient.py:24-40)."""
    torch = _torch()

    def dev(a):
        if torch.is_tensor(a):
            return a.to(device="cuda", dtype=torch.float64).contiguous().reshape(-1)
        return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64).reshape(-1)).to("cuda")
    f, a, v = dev(forward_wavefield), dev(adjoint_wavefield), dev(velocity)
    n = min(f.numel(), a.numel(), v.numel())
    g = torch.empty(n, dtype=torch.float64, device="cuda")
    call("stp_adjoint_gradient_f64", C.c_int64(n), ptr(f), ptr(a), ptr(v), C.c_double(dt), ptr(g), stream_handle())
    return g if torch.is_tensor(forward_wavefield) else g.cpu().numpy()
