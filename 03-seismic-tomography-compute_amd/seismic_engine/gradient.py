"""GradientComputer of the reference's Python inversion
(src/python/seismic_tomography/gradient.py:11-139) on the GPU: the same
methods and arguments, FP64 device tensors in and out (numpy arrays are
accepted and copied in; results stay on the device)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import call, ptr, stream_handle
from .rtm import compute_adjoint_gradient as _adjoint_gradient


def _torch():
    import torch
    return torch


def _dev(a):
    torch = _torch()
    if torch.is_tensor(a):
        return a.to(device="cuda", dtype=torch.float64).reshape(-1).contiguous()
    return torch.as_tensor(np.ascontiguousarray(np.asarray(a, np.float64).reshape(-1))).to("cuda")


class GradientComputer:
    def __init__(self, nx: int, ny: int, nz: int, smoothing_sigma: float = 2.0, regularization_lambda: float = 1e-4):
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.smoothing_sigma = float(smoothing_sigma)
        self.regularization_lambda = float(regularization_lambda)
        self._gradient = None
        self._preconditioner = None

    @property
    def n(self) -> int:
        return self.nx * self.ny * self.nz

    def _current(self, gradient):
        if gradient is None:
            gradient = self._gradient
        if gradient is None:
            raise ValueError("No gradient computed yet")
        return _dev(gradient)

    def compute_adjoint_gradient(self, forward_wavefield, adjoint_wavefield, velocity, dt: float):
        """-fwd * adj * 2 v dt^2 (gradient.py:24-40)."""
        g = _adjoint_gradient(_dev(forward_wavefield), _dev(adjoint_wavefield), _dev(velocity), dt)
        self._gradient = g
        return g

    def compute_frechet_gradient(self, residual, sensitivity_matrix):
        """A^T r (gradient.py:42-55; the reference's C++ branch calls
        SparseOps.dot_product instead, which cannot take a matrix): a
        DeviceCSR / CSRMatrix of this package, or anything with `.T @`."""
        from .sparse import CSRMatrix, DeviceCSR
        if isinstance(sensitivity_matrix, CSRMatrix):
            sensitivity_matrix = sensitivity_matrix.device()
        if isinstance(sensitivity_matrix, DeviceCSR):
            g = sensitivity_matrix.matvec(_dev(residual), transpose=True)
        else:
            g = _dev(sensitivity_matrix.T @ np.asarray(residual))
        self._gradient = g
        return g

    def apply_smoothing(self, gradient=None):
        """Gaussian smoothing, radius int(3 sigma), renormalised at the faces
        (SparseOps.smooth_gradient semantics, sparse_ops.cpp:115-158)."""
        from .sparse import SparseOps
        out = SparseOps.smooth_gradient(self._current(gradient), self.nx, self.ny, self.nz, self.smoothing_sigma)
        self._gradient = out
        return out

    def apply_preconditioning(self, gradient=None, preconditioner=None):
        g = self._current(gradient).clone()
        if preconditioner is not None:
            self._preconditioner = _dev(preconditioner)
        if self._preconditioner is not None:
            call("stp_gradient_precondition_f64", C.c_int64(g.numel()), ptr(self._preconditioner), ptr(g),
                 stream_handle())
        self._gradient = g
        return g

    def build_depth_preconditioner(self, water_depth_iz: int, dz: float = 200.0):
        p = _torch().empty(self.n, dtype=_torch().float64, device="cuda")
        call("stp_depth_preconditioner_f64", C.c_int64(self.nx), C.c_int64(self.ny), C.c_int64(self.nz),
             C.c_int64(int(water_depth_iz)), C.c_double(dz), ptr(p), stream_handle())
        self._preconditioner = p
        return p

    def apply_regularization(self, velocity, gradient=None):
        g = self._current(gradient).clone()
        v = _dev(velocity)
        call("stp_gradient_tikhonov_f64", C.c_int64(self.nx), C.c_int64(self.ny), C.c_int64(self.nz),
             C.c_double(self.regularization_lambda), ptr(v), ptr(g), stream_handle())
        self._gradient = g
        return g

    def normalize_gradient(self, gradient=None):
        g = self._current(gradient).clone()
        call("stp_gradient_normalize_f64", C.c_int64(g.numel()), ptr(g), None, stream_handle())
        self._gradient = g
        return g
