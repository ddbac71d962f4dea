"""ctypes wrappers of the FD oracle (oracle/fd_oracle.c).  Test infrastructure only."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_libs: dict = {}
_build = "check"


def select_build(kind: str) -> None:
    """"check" (default): liboracle.so, the parity checker; "timing":
    liboracle_timing.so, the same sources built with the reference's release
    flags (-ffast-math), for the CPU baseline timings only."""
    global _build
    if kind not in ("check", "timing"):
        raise ValueError(kind)
    _build = kind


class OrcCfg(C.Structure):
    _fields_ = [
        ("nx", C.c_int64), ("ny", C.c_int64), ("nz", C.c_int64),
        ("hw", C.c_int32), ("order", C.c_int32),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("dt", C.c_double),
        ("sponge_width", C.c_int32), ("sponge_damping", C.c_double),
        ("amplitude_limit", C.c_double), ("cfl_safety_factor", C.c_double), ("sponge_mode", C.c_int32),
    ]


class OrcResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("timesteps_completed", C.c_int32), ("completed_successfully", C.c_int32),
                ("total_nan_encountered", C.c_int64), ("total_cells_repaired", C.c_int64)]


def lib() -> C.CDLL:
    if _build not in _libs:
        p = HERE / ("liboracle.so" if _build == "check" else "liboracle_timing.so")
        if not p.exists():
            from . import build as _b
            _b.build_oracle()
        h = C.CDLL(str(p))
        h.orc_fd_stability_limit.restype = C.c_double
        h.orc_fd_sponge_factor.restype = C.c_double
        h.orc_count_nan_inf.restype = C.c_int64
        h.orc_repair_field_nan.restype = C.c_int64
        _libs[_build] = h
    return _libs[_build]


SPONGE_MODES = {"reference": 0, "cerjan": 1}


def cfg(n, hw=4, order=8, h=(10.0, 10.0, 10.0), dt=1e-3, sponge_width=20, sponge_damping=0.015,
        amplitude_limit=1e10, cfl_safety_factor=0.49, sponge="reference") -> OrcCfg:
    nx, ny, nz = n
    return OrcCfg(nx, ny, nz, hw, order, h[0], h[1], h[2], dt, sponge_width, sponge_damping, amplitude_limit,
                  cfl_safety_factor, SPONGE_MODES[sponge])


def _d(a):
    return a.ctypes.data_as(C.c_void_p)


def nh(c: OrcCfg):
    return (c.nx + 2 * c.hw, c.ny + 2 * c.hw, c.nz + 2 * c.hw)


def forward(c: OrcCfg, vel_h: np.ndarray, wavelet: np.ndarray, nt: int, src=None, rcv=None):
    """forward_propagate: returns (cur, prev, traces, result)."""
    shp = nh(c)
    vel_h = np.ascontiguousarray(vel_h, dtype=np.float64).reshape(shp)
    w = np.ascontiguousarray(wavelet, dtype=np.float64)
    if src is None:
        src = (c.nx // 2, c.ny // 2, c.nz // 2)
    rcv = np.ascontiguousarray(np.zeros(0, np.int64) if rcv is None else rcv, dtype=np.int64)
    tr = np.zeros((max(rcv.size, 1), nt), np.float64)
    cur = np.zeros(shp, np.float64)
    prev = np.zeros(shp, np.float64)
    res = OrcResult()
    lib().orc_fd_forward(C.byref(c), _d(vel_h), _d(w), C.c_int32(w.size), C.c_int32(nt), C.c_int64(src[0]),
                         C.c_int64(src[1]), C.c_int64(src[2]), _d(rcv), C.c_int32(rcv.size), _d(tr), _d(cur),
                         _d(prev), C.byref(res))
    return cur, prev, tr[: rcv.size], res


def run(c: OrcCfg, vel_h, cur, prev, nsteps, src_idx=(), src_w=None, rcv=None):
    """Step from a given state (no reset / CFL check); arrays updated in place.
    Returns (status, traces)."""
    shp = nh(c)
    assert cur.shape == shp and prev.shape == shp and cur.dtype == np.float64
    vel_h = np.ascontiguousarray(vel_h, dtype=np.float64).reshape(shp)
    src_idx = np.ascontiguousarray(np.asarray(src_idx, dtype=np.int64).reshape(-1))
    nsrc = src_idx.size
    if src_w is None:
        src_w = np.zeros((max(nsrc, 1), 0))
    src_w = np.ascontiguousarray(np.asarray(src_w, dtype=np.float64).reshape(max(nsrc, 1), -1))
    nw = src_w.shape[1] if nsrc else 0
    rcv = np.ascontiguousarray(np.zeros(0, np.int64) if rcv is None else rcv, dtype=np.int64)
    tr = np.zeros((max(rcv.size, 1), max(nsteps, 1)), np.float64)
    s = lib().orc_fd_run(C.byref(c), _d(vel_h), _d(cur), _d(prev), C.c_int32(nsteps), _d(src_idx), _d(src_w),
                         C.c_int32(nsrc), C.c_int32(nw), _d(rcv), C.c_int32(rcv.size), _d(tr))
    return s, tr[: rcv.size, :nsteps]


def stencil(c: OrcCfg, cur, prev, vel_h):
    nxt = np.zeros_like(cur)
    s = lib().orc_fd_stencil(C.byref(c), _d(cur), _d(prev), _d(nxt), _d(np.ascontiguousarray(vel_h, np.float64)))
    return s, nxt


def sponge_factors(c: OrcCfg) -> np.ndarray:
    shp = nh(c)
    out = np.empty(shp)
    f = lib().orc_fd_sponge_factor
    f.argtypes = [C.POINTER(OrcCfg), C.c_int64, C.c_int64, C.c_int64]
    for ix in range(shp[0]):
        for iy in range(shp[1]):
            for iz in range(shp[2]):
                out[ix, iy, iz] = f(C.byref(c), ix, iy, iz)
    return out


def weights(order: int) -> np.ndarray:
    a = np.zeros(5)
    R = lib().orc_fd_weights(order, _d(a))
    return a[: R + 1]


def stability_limit(order: int) -> float:
    return lib().orc_fd_stability_limit(order)


def verify_cfl(c: OrcCfg, vel_h) -> bool:
    return bool(lib().orc_fd_verify_cfl(C.byref(c), _d(np.ascontiguousarray(vel_h, np.float64))))


def ricker(f, dt, t0, n) -> np.ndarray:
    out = np.empty(n)
    lib().orc_ricker(C.c_double(f), C.c_double(dt), C.c_double(t0), C.c_int32(n), _d(out))
    return out


def repair_field_nan(data: np.ndarray, nx, ny, nz, hw=1, replace_value=0.0) -> np.ndarray:
    d = np.ascontiguousarray(data, dtype=np.float64).copy()
    lib().orc_repair_field_nan(_d(d), C.c_int64(d.size), C.c_int64(nx), C.c_int64(ny), C.c_int64(nz),
                               C.c_int(hw), C.c_double(replace_value))
    return d
