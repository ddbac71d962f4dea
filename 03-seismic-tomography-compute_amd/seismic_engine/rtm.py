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


# ------------------------------------------------------------------ mmap snapshot file
# Version 2 of the reference's file (mmap_writer.h:28-39): the 120-byte
# header (magic, header_size, file_size_bytes, frame_count, write_offset,
# ring_start, ring_capacity, 8 reserved words) zero-padded to header_size =
# 128, then [i64 size][frame] records from 128 to write_offset.  The
# reference's C++ writer starts frames at 120 and its Python reader at 128;
# here both sides use header_size.  A record that does not fit behind
# write_offset restarts the ring at header_size (ring_start counts restarts)
# and frame_count is the number of records between header_size and
# write_offset.
MMAP_FILE_MAGIC = 0x52544D534E4150
MMAP_HEADER_BYTES = 128
_FHDR = struct.Struct("<qqqqqqq")


class MmapSnapshotWriter:
    """MmapSnapshotWriter (mmap_writer.h:41-121) over Python's mmap."""

    def __init__(self):
        self._mm = None
        self._fh = None
        self.max_bytes = 0
        self.frames_written = 0
        self.frames_dropped = 0
        self.bytes_written = 0
        self.cycles = 0
        self._seq = 0

    def open(self, filename: str, max_bytes: int = 4 * 1024 ** 3) -> bool:
        import mmap
        self.close()
        if max_bytes < MMAP_HEADER_BYTES + 8:
            return False
        self._fh = open(filename, "w+b")
        self._fh.truncate(max_bytes)
        self._mm = mmap.mmap(self._fh.fileno(), max_bytes)
        self.max_bytes = int(max_bytes)
        self._mm[:MMAP_HEADER_BYTES] = bytes(MMAP_HEADER_BYTES)
        _FHDR.pack_into(self._mm, 0, MMAP_FILE_MAGIC, MMAP_HEADER_BYTES, self.max_bytes, 0, MMAP_HEADER_BYTES, 0, 0)
        self.cycles += 1
        return True

    def is_open(self) -> bool:
        return self._mm is not None

    def _field(self, k: int) -> int:
        return struct.unpack_from("<q", self._mm, 8 * k)[0]

    def _set(self, k: int, v: int) -> None:
        struct.pack_into("<q", self._mm, 8 * k, int(v))

    def frame_count(self) -> int:
        return self._field(3)

    def write_offset(self) -> int:
        return self._field(4)

    def write_frame(self, frame: bytes) -> bool:
        n = len(frame)
        if self._mm is None or n <= 0 or n > self.max_bytes - MMAP_HEADER_BYTES - 8:
            self.frames_dropped += 1
            return False
        off = self._field(4)
        if off + 8 + n > self.max_bytes:
            self._set(3, 0)
            self._set(4, MMAP_HEADER_BYTES)
            self._set(5, self._field(5) + 1)
            self.cycles += 1
            off = MMAP_HEADER_BYTES
        struct.pack_into("<q", self._mm, off, n)
        self._mm[off + 8:off + 8 + n] = frame
        self._set(4, off + 8 + n)
        self._set(3, self._field(3) + 1)
        self.frames_written += 1
        self.bytes_written += 8 + n
        return True

    def write_slice(self, sl: CapturedSlice) -> bool:
        self._seq += 1
        return self.write_frame(encode_slice_frame(sl, self._seq - 1))

    def close(self) -> None:
        if self._mm is not None:
            self._mm.flush()
            self._mm.close()
            self._fh.close()
        self._mm = None
        self._fh = None


def read_snapshot_file(filename: str) -> list[bytes]:
    """The records of a snapshot file, oldest first (read_frame,
    mmap_writer.cpp:230-290); [] when the magic or header is wrong."""
    with open(filename, "rb") as f:
        buf = f.read()
    if len(buf) < MMAP_HEADER_BYTES:
        return []
    magic, hsize, _, _, end, _, _ = _FHDR.unpack_from(buf, 0)
    if magic != MMAP_FILE_MAGIC or not 120 <= hsize <= len(buf):
        return []
    end = min(end, len(buf))
    out, off = [], hsize
    while off + 8 <= end:
        n = struct.unpack_from("<q", buf, off)[0]
        off += 8
        if n <= 0 or n > end - off:
            break
        out.append(buf[off:off + n])
        off += n
    return out


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
