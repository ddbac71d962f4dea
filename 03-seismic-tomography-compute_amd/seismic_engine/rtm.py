"""RTM imaging types of the reference's module (bindings.cpp:303-371) and an
RTMEngine with the reference's methods (src/cpp/rtm/rtm_engine.h:31-96):
FP64 image and illumination kept on the GPU; slices for live viewing are
gathered on the GPU (stp_rtm_extract_slice_f64) and handed to the slice
callback.  The frame codec writes and reads the slice frames of the
reference's snapshot buffer (src/cpp/rtm/snapshot_buffer.h:18-37,
snapshot_buffer.cpp:100-227)."""
from __future__ import annotations

import ctypes as C
import enum
import struct
import time
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
class CapturedSlice:                    # rtm_engine.h:63-73
    timestep: int
    axis: SliceAxis
    index: int
    dim0: int
    dim1: int
    forward_slice: np.ndarray
    adjoint_slice: np.ndarray
    ic_slice: np.ndarray
    capture_timestamp: float = 0.0


# ------------------------------------------------------------------ slice frames
# The reference's SliceFrameHeader is 124 packed bytes but frames reserve 128
# (SLICE_HEADER_SIZE) and copy 128 bytes from the struct, so bytes 124-127 of
# its frames are whatever followed the struct in memory.  These frames use the
# same field layout with those four bytes zero; payload offsets are the
# reference's (forward at 128, then adjoint, then image condition, FP64).
SLICE_MAGIC = 0x534C494345303031
SLICE_HEADER_SIZE = 128
_HDR = struct.Struct("<QqiqqqqqqqqqqdQII")       # 124 bytes, packed


def slice_checksum(payload: bytes) -> int:
    """compute_checksum (snapshot_buffer.cpp:206-227): xor of little-endian
    32-bit words, each followed by a rotate left by 13; a trailing partial
    word is xor-ed in without rotation."""
    n32 = len(payload) // 4
    words = np.frombuffer(payload[: 4 * n32], dtype="<u4")
    s = 0
    for w in words.tolist():
        s ^= w
        s = ((s << 13) | (s >> 19)) & 0xFFFFFFFF
    rest = payload[4 * n32:]
    if rest:
        s ^= int.from_bytes(rest, "little")
    return s


def encode_slice_frame(sl: CapturedSlice, sequence_id: int = 0) -> bytes:
    """serialize_slice (snapshot_buffer.cpp:100-149)."""
    parts = [np.ascontiguousarray(np.asarray(a, dtype="<f8")).tobytes()
             for a in (sl.forward_slice, sl.adjoint_slice, sl.ic_slice)]
    fb, ab, ib = (len(p) for p in parts)
    payload = b"".join(parts)
    fo = SLICE_HEADER_SIZE
    hdr = _HDR.pack(SLICE_MAGIC, int(sl.timestep), int(sl.axis), int(sl.index), int(sl.dim0), int(sl.dim1),
                    SLICE_HEADER_SIZE + len(payload), fo, fo + fb, fo + fb + ab, fb, ab, ib,
                    float(sl.capture_timestamp), int(sequence_id), slice_checksum(payload), 0)
    return hdr + bytes(SLICE_HEADER_SIZE - _HDR.size) + payload


def decode_slice_frame(buf: bytes):
    """deserialize_slice (snapshot_buffer.cpp:151-204): None unless the magic,
    the sizes and the checksum agree; else (CapturedSlice, sequence_id)."""
    if len(buf) < SLICE_HEADER_SIZE:
        return None
    (magic, ts, axis, index, d0, d1, size, fo, ao, io, fb, ab, ib, stamp, seq, csum,
     _) = _HDR.unpack_from(buf, 0)
    if magic != SLICE_MAGIC or size > len(buf) or size < SLICE_HEADER_SIZE:
        return None
    if slice_checksum(bytes(buf[SLICE_HEADER_SIZE:size])) != csum:
        return None
    if min(fb, ab, ib) < 0 or fo + fb > size or ao + ab > size or io + ib > size:
        return None

    def arr(o, b):
        return np.frombuffer(bytes(buf[o:o + b]), dtype="<f8").copy()
    return CapturedSlice(ts, SliceAxis(axis), index, d0, d1, arr(fo, fb), arr(ao, ab), arr(io, ib), stamp), seq


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
        self.slice_cb = None
        self._captured = 0

    def set_slice_callback(self, cb) -> None:
        self.slice_cb = cb

    def total_captured_slices(self) -> int:
        return self._captured

    def extract_slices(self, forward_wavefield, adjoint_wavefield, image_condition, ny: int, nz: int,
                       timestep: int) -> None:
        """rtm_engine.cpp:59-81: for each slice spec (a negative index means
        the middle of that axis, VOLUME is skipped) gather the plane of the
        forward and adjoint with-halo fields (y / z strides ny, nz) and of
        the image condition (the engine's image when None) on the GPU and
        pass a CapturedSlice to the callback."""
        if self.slice_cb is None or not self.slice_specs:
            return
        torch = _torch()

        def dev(a):
            if torch.is_tensor(a):
                return a.to(device="cuda", dtype=torch.float64).contiguous().reshape(-1)
            return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64).reshape(-1)).to("cuda")
        f, a = dev(forward_wavefield), dev(adjoint_wavefield)
        ic = self.image if image_condition is None else dev(image_condition)
        ext = (self.nx, self.ny, self.nz)
        for spec in self.slice_specs:
            axis = SliceAxis(spec.axis)
            if axis == SliceAxis.VOLUME:
                continue
            idx = int(spec.index) if spec.index >= 0 else ext[int(axis)] // 2
            d0 = self.ny if axis == SliceAxis.X else self.nx
            d1 = self.ny if axis == SliceAxis.Z else self.nz
            out = torch.empty(3, d0 * d1, dtype=torch.float64, device="cuda")
            call("stp_rtm_extract_slice_f64", int(axis), C.c_int64(idx), C.c_int64(self.nx), C.c_int64(self.ny),
                 C.c_int64(self.nz), self.halo_width, C.c_int64(ny), C.c_int64(nz), ptr(f), ptr(a), ptr(ic),
                 ptr(out[0]), ptr(out[1]), ptr(out[2]), stream_handle())
            h = out.cpu().numpy()
            self.slice_cb(CapturedSlice(int(timestep), axis, idx, d0, d1, h[0], h[1], h[2], time.time()))
            self._captured += 1

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
