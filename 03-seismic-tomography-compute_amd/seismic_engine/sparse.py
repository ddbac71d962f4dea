"""Sparse matrices and solvers on the GPU, behind the reference's pybind11
names (src/cpp/pybind/bindings.cpp:52-193, 243-253).

CSRMatrix keeps the canonical FP64/int64 CSR on the device (what row_ptr(),
col_ind() and values() return) and derives the FP32/int32 device matrix plus
its transpose (stp_csr) for SpMV and the solvers on first use.  Arrays may be
numpy (copied, results returned as numpy float64 like the reference) or
torch tensors (kept on the device).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import SolveInfo, call, lib, ptr, stream_handle


def _torch():
    import torch
    return torch


def _dev():
    return _torch().device("cuda", _torch().cuda.current_device())


def _as_dev(a, dtype):
    torch = _torch()
    if torch.is_tensor(a):
        return a.to(device=_dev(), dtype=dtype).contiguous().reshape(-1)
    return torch.as_tensor(np.ascontiguousarray(np.asarray(a)).reshape(-1)).to(device=_dev(), dtype=dtype)


def _out(t, like):
    torch = _torch()
    return t if torch.is_tensor(like) else t.cpu().numpy()


class DeviceCSR:
    """FP32-valued device CSR with an explicit transpose (stp_csr)."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle
        r, c, n = C.c_int64(), C.c_int64(), C.c_int64()
        call("stp_csr_shape", self._h, C.byref(r), C.byref(c), C.byref(n))
        self.rows, self.cols, self.nnz = r.value, c.value, n.value

    @classmethod
    def from_arrays(cls, row_ptr, col, val):
        torch = _torch()
        rp = _as_dev(row_ptr, torch.int32)
        ci = _as_dev(col, torch.int32)
        v = _as_dev(val, torch.float32)
        rows = rp.numel() - 1
        cols = int(ci.max().item()) + 1 if ci.numel() else 0
        return cls.from_device(rows, cols, rp, ci, v)

    @classmethod
    def from_device(cls, rows, cols, rp32, ci32, v32, nnz=None):
        h = C.c_void_p()
        nnz = ci32.numel() if nnz is None else nnz
        call("stp_csr_create", C.c_int64(rows), C.c_int64(cols), C.c_int64(nnz), ptr(rp32), ptr(ci32),
             ptr(v32), C.byref(h), stream_handle())
        return cls(h)

    @classmethod
    def from_rays(cls, pts, n, h: float = 1.0):
        """Exact straight-ray matrix of rays pts[r] = (p0, p1) through an
        n = (nx, ny, nz) grid of cubic cells of size h."""
        torch = _torch()
        p = _as_dev(pts, torch.float64)
        nn = np.ascontiguousarray(np.asarray(n, dtype=np.int64))
        hh = C.c_void_p()
        call("stp_csr_from_rays", C.c_int64(p.numel() // 6), nn.ctypes.data, C.c_double(h), ptr(p), C.byref(hh),
             stream_handle())
        return cls(hh)

    def set_grid(self, n) -> None:
        """Declare the columns as the cells of an n = (nx, ny, nz) grid (None
        clears it): CGLS then runs in the Morton solver layout."""
        nx, ny, nz = (0, 0, 0) if n is None else (int(k) for k in n)
        call("stp_csr_set_grid", self._h, C.c_int64(nx), C.c_int64(ny), C.c_int64(nz))

    def prepare_layout(self) -> None:
        """Build the solver layout now instead of on the first CGLS solve."""
        call("stp_csr_prepare_layout", self._h, stream_handle())

    def staging(self) -> tuple[int, int]:
        """(rows per tile of the staged A^T, length of the tiles' lists of
        distinct rows); rows per tile 0 when A^T is not staged, -1 before the
        solver layout is built (stp_csr_staging)."""
        t, n = C.c_int32(), C.c_int64()
        call("stp_csr_staging", self._h, C.byref(t), C.byref(n))
        return t.value, n.value

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().stp_csr_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def export(self, transpose: bool = False):
        torch = _torch()
        n = self.cols if transpose else self.rows
        rp = torch.empty(n + 1, dtype=torch.int32, device=_dev())
        ci = torch.empty(max(self.nnz, 1), dtype=torch.int32, device=_dev())
        v = torch.empty(max(self.nnz, 1), dtype=torch.float32, device=_dev())
        call("stp_csr_export", self._h, int(transpose), ptr(rp), ptr(ci), ptr(v), stream_handle())
        return rp, ci[: self.nnz], v[: self.nnz]

    def matvec(self, x, transpose: bool = False):
        """y = A x or A^T x.  FP32 x -> FP32 accumulation; FP64 x -> FP64."""
        torch = _torch()
        n_out = self.cols if transpose else self.rows
        if x.dtype == torch.float64:
            y = torch.empty(n_out, dtype=torch.float64, device=x.device)
            call("stp_csr_spmv_f64", self._h, int(transpose), ptr(x), ptr(y), stream_handle())
        else:
            y = torch.empty(n_out, dtype=torch.float32, device=x.device)
            call("stp_csr_spmv", self._h, int(transpose), ptr(x.float().contiguous()), ptr(y), stream_handle())
        return y

    def cg(self, b, max_iter=1000, tol=1e-8, jacobi=False, history=False):
        torch = _torch()
        bb = _as_dev(b, torch.float64)
        x = torch.empty(self.cols, dtype=torch.float64, device=bb.device)
        hist = torch.zeros(max(max_iter, 1), dtype=torch.float64, device=bb.device) if history else None
        info = SolveInfo()
        call("stp_cg", self._h, ptr(bb), int(max_iter), C.c_double(tol), int(jacobi), ptr(x), ptr(hist),
             C.byref(info), stream_handle())
        return (x, info, hist[: info.iterations] if history else None)

    def cgls(self, b, lam: float, tol: float = 1e-6, max_iter: int = 10000, precond: str = "none",
             network=None, history=False, fallback: bool = True):
        """CGLS on (A^T A + lam I) x = A^T b.  precond "none", "jacobi" or
        "learned" (a prepared LearnedPreconditioner in `network`); a learned
        solve that does not reach tol in max_iter is redone with Jacobi when
        `fallback` (info.fallback tells)."""
        torch = _torch()
        kind = {"none": 0, "jacobi": 1, "learned": 2}[precond]
        bb = _as_dev(b, torch.float64)
        x = torch.empty(self.cols, dtype=torch.float64, device=bb.device)
        hist = torch.zeros(max(max_iter, 1), dtype=torch.float64, device=bb.device) if history else None
        info = SolveInfo()
        hnet = network.handle if network is not None else None
        call("stp_cgls", self._h, ptr(bb), C.c_double(lam), C.c_double(tol), int(max_iter), kind, hnet, ptr(x),
             ptr(hist), C.byref(info), stream_handle())
        info.fallback = False
        if kind == 2 and fallback and not info.converged:
            # Jacobi-CGLS is the fallback of the learned preconditioner
            x, info, hist = self.cgls(b, lam, tol, max_iter, "jacobi", history=history)
            info.fallback = True
            return x, info, hist
        return x, info, (hist[: info.iterations] if history else None)


class CSRMatrix:
    """Drop-in for the reference CSRMatrix binding (bindings.cpp:52-138)."""

    def __init__(self, rows: int = 0, cols: int = 0, nnz: int = 0):
        torch = _torch()
        self._rows, self._cols = int(rows), int(cols)
        self._rp = torch.zeros(self._rows + 1, dtype=torch.int64, device=_dev())
        self._ci = torch.zeros(int(nnz), dtype=torch.int64, device=_dev())
        self._val = torch.zeros(int(nnz), dtype=torch.float64, device=_dev())
        self._dev = None
        self.block_size = 8

    # ---- construction
    def _assemble(self, rows, cols, r, c, v):
        torch = _torch()
        n = r.numel()
        rp = torch.empty(rows + 1, dtype=torch.int64, device=_dev())
        co = torch.empty(max(n, 1), dtype=torch.int64, device=_dev())
        vo = torch.empty(max(n, 1), dtype=torch.float64, device=_dev())
        nnz = C.c_int64()
        call("stp_csr_assemble", C.c_int64(rows), C.c_int64(cols), C.c_int64(n), ptr(r), ptr(c), ptr(v), ptr(rp),
             ptr(co), ptr(vo), C.byref(nnz), stream_handle())
        self._rows, self._cols = int(rows), int(cols)
        self._rp, self._ci, self._val = rp, co[: nnz.value].clone(), vo[: nnz.value].clone()
        self._dev = None

    def build_from_triplets(self, rows, cols, row_idx, col_idx, values):
        torch = _torch()
        r = _as_dev(row_idx, torch.int64)
        c = _as_dev(col_idx, torch.int64)
        v = _as_dev(values, torch.float64)
        if not (r.numel() == c.numel() == v.numel()):
            raise ValueError("row_idx, col_idx and values must have the same length")
        self._assemble(int(rows), int(cols), r, c, v)

    def build_fd_stencil_7point(self, nx, ny, nz, velocity, dx, dy, dz, dt):
        torch = _torch()
        n = int(nx) * int(ny) * int(nz)
        vel = _as_dev(velocity, torch.float64)
        if vel.numel() < n:
            raise ValueError("velocity has fewer than nx*ny*nz values")
        r = torch.empty(7 * n, dtype=torch.int64, device=_dev())
        c = torch.empty_like(r)
        v = torch.empty(7 * n, dtype=torch.float64, device=_dev())
        call("stp_csr64_fd7_triplets", C.c_int64(nx), C.c_int64(ny), C.c_int64(nz), ptr(vel), C.c_double(dx),
             C.c_double(dy), C.c_double(dz), C.c_double(dt), ptr(r), ptr(c), ptr(v), stream_handle())
        self._assemble(n, n, r, c, v)

    # ---- device matrix
    def device(self) -> DeviceCSR:
        torch = _torch()
        if self._dev is None:
            nnz = self._ci.numel()
            rp32 = torch.empty(self._rows + 1, dtype=torch.int32, device=_dev())
            ci32 = torch.empty(max(nnz, 1), dtype=torch.int32, device=_dev())
            v32 = torch.empty(max(nnz, 1), dtype=torch.float32, device=_dev())
            call("stp_csr64_to_f32", C.c_int64(self._rows), C.c_int64(nnz), ptr(self._rp), ptr(self._ci),
                 ptr(self._val), ptr(rp32), ptr(ci32), ptr(v32), stream_handle())
            self._dev = DeviceCSR.from_device(self._rows, self._cols, rp32, ci32, v32, nnz)
            # FP64 values for the reference's all-FP64 CG (SparseOps.conjugate_gradient)
            call("stp_csr_set_values_f64", self._dev._h, ptr(self._val), stream_handle())
        return self._dev

    # ---- reference methods
    def spmv(self, x):
        torch = _torch()
        xv = _as_dev(x, torch.float64)
        if xv.numel() < self._cols:
            raise ValueError("x is shorter than cols()")
        return _out(self.device().matvec(xv[: self._cols]), x)

    def spmv_transpose(self, x):
        torch = _torch()
        xv = _as_dev(x, torch.float64)
        if xv.numel() < self._rows:
            raise ValueError("x is shorter than rows()")
        return _out(self.device().matvec(xv[: self._rows], transpose=True), x)

    def frobenius_norm(self) -> float:
        out = C.c_double()
        call("stp_dot_f64", C.c_int64(self._val.numel()), ptr(self._val), ptr(self._val), C.byref(out),
             stream_handle())
        return float(np.sqrt(out.value))

    def scale(self, alpha: float) -> None:
        call("stp_csr64_scale", C.c_int64(self._val.numel()), ptr(self._val), C.c_double(alpha), stream_handle())
        self._dev = None

    def add_diagonal(self, alpha: float) -> None:
        call("stp_csr64_add_diagonal", C.c_int64(self._rows), ptr(self._rp), ptr(self._ci), ptr(self._val),
             C.c_double(alpha), stream_handle())
        self._dev = None

    def optimize_cache_blocking(self, block_size: int = 8) -> None:
        """x86 cache-blocking hint of the reference (csr_matrix.cpp:195-210);
        the GPU SpMV has its own row mapping, so only the value is kept."""
        self.block_size = int(block_size)

    def rows(self) -> int:
        return self._rows

    def cols(self) -> int:
        return self._cols

    def nnz(self) -> int:
        return int(self._ci.numel())

    def row_ptr(self):
        return self._rp.cpu().numpy()

    def col_ind(self):
        return self._ci.cpu().numpy()

    def values(self):
        return self._val.cpu().numpy()

    def copy(self) -> "CSRMatrix":
        m = CSRMatrix()
        m._rows, m._cols = self._rows, self._cols
        m._rp, m._ci, m._val = self._rp.clone(), self._ci.clone(), self._val.clone()
        return m


def compute_fd_stencil(nx, ny, nz, velocity, dx, dy, dz, dt) -> CSRMatrix:
    """bindings.cpp:243-253: the 7-point FD step as a CSR matrix."""
    m = CSRMatrix()
    m.build_fd_stencil_7point(nx, ny, nz, velocity, dx, dy, dz, dt)
    return m


class SparseOps:
    """Drop-in for the reference SparseOps binding (bindings.cpp:140-193)."""

    @staticmethod
    def conjugate_gradient(A: CSRMatrix, b, max_iter: int = 1000, tolerance: float = 1e-8):
        x, _, _ = A.device().cg(b, max_iter, tolerance, jacobi=False)
        return _out(x, b)

    @staticmethod
    def jacobi_preconditioned_cg(A: CSRMatrix, b, max_iter: int = 1000, tolerance: float = 1e-8):
        x, _, _ = A.device().cg(b, max_iter, tolerance, jacobi=True)
        return _out(x, b)

    @staticmethod
    def smooth_gradient(gradient, nx: int, ny: int, nz: int, sigma: float):
        """Gaussian smoothing with radius int(3 sigma) and in-range
        renormalisation (sparse_ops.cpp:115-158), FP64 on the GPU as three
        separable passes; a gradient shorter than nx*ny*nz gives zeros."""
        torch = _torch()
        n = int(nx) * int(ny) * int(nz)
        g = _as_dev(gradient, torch.float64)
        if g.numel() < n:
            return _out(torch.zeros(n, dtype=torch.float64, device=g.device), gradient)
        g = g[:n].contiguous()
        out = torch.empty_like(g)
        work = torch.empty_like(g)
        call("stp_smooth_gradient_f64", C.c_int64(nx), C.c_int64(ny), C.c_int64(nz), C.c_double(sigma), ptr(g),
             ptr(out), ptr(work), stream_handle())
        return _out(out, gradient)

    @staticmethod
    def l2_norm(a) -> float:
        return float(np.sqrt(SparseOps.dot_product(a, a)))

    @staticmethod
    def dot_product(a, b) -> float:
        if isinstance(a, CSRMatrix) or isinstance(b, CSRMatrix):
            raise TypeError("dot_product(a, b) takes two 1-D arrays")
        torch = _torch()
        av = _as_dev(a, torch.float64)
        bv = _as_dev(b, torch.float64)
        n = min(av.numel(), bv.numel())
        out = C.c_double()
        call("stp_dot_f64", C.c_int64(n), ptr(av), ptr(bv), C.byref(out), stream_handle())
        return out.value

    @staticmethod
    def build_hessian_approx(forward_operator: CSRMatrix, regularization: float) -> CSRMatrix:
        """A + lambda I on the first diagonal entry of each row (sparse_ops.cpp:204-210)."""
        h = forward_operator.copy()
        h.add_diagonal(regularization)
        return h

    @staticmethod
    def cgls(A, b, regularization: float = 1e-2, tolerance: float = 1e-6, max_iter: int = 10000,
             preconditioner: str = "jacobi", network=None):
        """x minimising |A x - b|^2 + lambda |x|^2 (normal equations
        (A^T A + lambda I) x = A^T b) on the GPU; returns (x, info dict)."""
        dev = A.device() if isinstance(A, CSRMatrix) else A
        x, info, _ = dev.cgls(b, regularization, tolerance, max_iter, preconditioner, network)
        return _out(x, b), info.as_dict()
