"""ctypes wrappers of the sparse oracle (oracle/sparse_oracle.c).  Test infrastructure only."""
from __future__ import annotations

import ctypes as C

import numpy as np

from .fd import lib as _lib


def _d(a):
    return a.ctypes.data_as(C.c_void_p)


def csr_from_triplets(rows, cols, r, c, v):
    r = np.ascontiguousarray(r, np.int64); c = np.ascontiguousarray(c, np.int64)
    v = np.ascontiguousarray(v, np.float64)
    rp = np.zeros(rows + 1, np.int64)
    ci = np.zeros(max(r.size, 1), np.int64)
    val = np.zeros(max(r.size, 1), np.float64)
    f = _lib().orc_csr_from_triplets
    f.restype = C.c_int64
    nnz = f(C.c_int64(rows), C.c_int64(cols), C.c_int64(r.size), _d(r), _d(c), _d(v), _d(rp), _d(ci), _d(val))
    return rp, ci[:nnz].copy(), val[:nnz].copy()


def spmv(rp, ci, val, x):
    rows = rp.size - 1
    y = np.zeros(rows)
    _lib().orc_csr_spmv(C.c_int64(rows), _d(np.ascontiguousarray(rp, np.int64)), _d(np.ascontiguousarray(ci, np.int64)),
                        _d(np.ascontiguousarray(val, np.float64)), _d(np.ascontiguousarray(x, np.float64)), _d(y))
    return y


def spmv_t(rp, ci, val, x, cols):
    rows = rp.size - 1
    y = np.zeros(cols)
    _lib().orc_csr_spmv_t(C.c_int64(rows), C.c_int64(cols), _d(np.ascontiguousarray(rp, np.int64)),
                          _d(np.ascontiguousarray(ci, np.int64)), _d(np.ascontiguousarray(val, np.float64)),
                          _d(np.ascontiguousarray(x, np.float64)), _d(y))
    return y


def _solver(name, rp, ci, val, b, max_iter, tol):
    n = rp.size - 1
    x = np.zeros(n)
    rr = C.c_double()
    it = getattr(_lib(), name)(C.c_int64(n), _d(np.ascontiguousarray(rp, np.int64)),
                               _d(np.ascontiguousarray(ci, np.int64)), _d(np.ascontiguousarray(val, np.float64)),
                               _d(np.ascontiguousarray(b, np.float64)), C.c_int(max_iter), C.c_double(tol), _d(x),
                               C.byref(rr))
    return x, it, rr.value


def cg(rp, ci, val, b, max_iter=1000, tol=1e-8):
    return _solver("orc_cg", rp, ci, val, b, max_iter, tol)


def jacobi_pcg(rp, ci, val, b, max_iter=1000, tol=1e-8):
    return _solver("orc_jacobi_pcg", rp, ci, val, b, max_iter, tol)


def cgls(rp, ci, val, b, cols, lam, tol=1e-6, max_iter=10000, precond=0):
    rows = rp.size - 1
    x = np.zeros(cols)
    rr = C.c_double()
    hist = np.zeros(max_iter)
    it = _lib().orc_cgls(C.c_int64(rows), C.c_int64(cols), _d(np.ascontiguousarray(rp, np.int64)),
                         _d(np.ascontiguousarray(ci, np.int64)), _d(np.ascontiguousarray(val, np.float64)),
                         _d(np.ascontiguousarray(b, np.float64)), C.c_double(lam), C.c_double(tol),
                         C.c_int(max_iter), C.c_int(precond), _d(x), C.byref(rr), _d(hist))
    return x, it, rr.value, hist[:it]


def siddon(n, h, p0, p1, cap=4096):
    nn = np.ascontiguousarray(n, np.int64)
    a = np.ascontiguousarray(p0, np.float64); b = np.ascontiguousarray(p1, np.float64)
    cells = np.zeros(cap, np.int64)
    lens = np.zeros(cap)
    f = _lib().orc_siddon
    f.restype = C.c_int64
    m = f(_d(nn), C.c_double(h), _d(a), _d(b), _d(cells), _d(lens), C.c_int64(cap))
    if m < 0:
        raise ValueError("cap too small")
    return cells[:m].copy(), lens[:m].copy()


def ray_matrix(n, h, pts):
    """CSR (int64 row_ptr, int64 cols sorted per row, FP64 lengths) of the rays."""
    rows = []
    rp = [0]
    cols = []
    vals = []
    for p in pts:
        c, l = siddon(n, h, p[0], p[1])
        o = np.argsort(c, kind="stable")
        cols.append(c[o]); vals.append(l[o])
        rp.append(rp[-1] + c.size)
    return np.asarray(rp, np.int64), np.concatenate(cols), np.concatenate(vals)


def smooth_gradient(g, nx, ny, nz, sigma):
    """SparseOps::smooth_gradient restated (direct 3-D sum)."""
    g = np.ascontiguousarray(g, np.float64).reshape(-1)
    out = np.zeros(nx * ny * nz)
    if g.size < out.size:
        return out
    _lib().orc_smooth_gradient(C.c_int64(nx), C.c_int64(ny), C.c_int64(nz), C.c_double(sigma), _d(g), _d(out))
    return out
