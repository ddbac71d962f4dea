"""ctypes wrappers of the reverse-time oracle (oracle/adjoint_oracle.c).  Test infrastructure only."""
from __future__ import annotations

import ctypes as C

import numpy as np

from .fd import OrcCfg, _d, lib, nh


def adjoint(c: OrcCfg, vel_h, wavelet, nt, src_idx, rcv, inj=None, scale_c=True, snap_lap=True, ill=False):
    shp = nh(c)
    vel_h = np.ascontiguousarray(vel_h, dtype=np.float64).reshape(shp)
    w = np.ascontiguousarray(wavelet, dtype=np.float64)
    rcv = np.ascontiguousarray(rcv, dtype=np.int64)
    acc = np.zeros(c.nx * c.ny * c.nz)
    il = np.zeros_like(acc) if ill else None
    tr = np.zeros((rcv.size, nt))
    injv = None if inj is None else np.ascontiguousarray(inj, dtype=np.float64).reshape(rcv.size, nt)
    f = lib().orc_fd_adjoint
    f.restype = C.c_int
    rc = f(C.byref(c), _d(vel_h), _d(w), C.c_int32(w.size), C.c_int32(nt), C.c_int64(src_idx), _d(rcv),
           C.c_int32(rcv.size), None if injv is None else _d(injv), C.c_int32(int(scale_c)), C.c_int32(int(snap_lap)),
           _d(acc), None if il is None else _d(il), _d(tr))
    assert rc == 0
    n = (c.nx, c.ny, c.nz)
    return acc.reshape(n), (il.reshape(n) if il is not None else None), tr


def misfit(c: OrcCfg, vel_h, wavelet, nt, src_idx, rcv, obs) -> float:
    vel_h = np.ascontiguousarray(vel_h, dtype=np.float64)
    w = np.ascontiguousarray(wavelet, dtype=np.float64)
    rcv = np.ascontiguousarray(rcv, dtype=np.int64)
    o = np.ascontiguousarray(obs, dtype=np.float64)
    f = lib().orc_fd_misfit
    f.restype = C.c_double
    return f(C.byref(c), _d(vel_h), _d(w), C.c_int32(w.size), C.c_int32(nt), C.c_int64(src_idx), _d(rcv),
             C.c_int32(rcv.size), _d(o))


def laplacian(c: OrcCfg, u):
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros_like(u)
    lib().orc_fd_laplacian(C.byref(c), _d(u), _d(out))
    return out
