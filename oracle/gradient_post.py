"""CPU oracle of the gradient post-processing and of the travel-time
inversion loop.  TEST INFRASTRUCTURE ONLY.

Restates GradientComputer (src/python/seismic_tomography/gradient.py:77-139:
preconditioning, depth weights, Laplacian Tikhonov term, max-abs
normalisation) in numpy, and AdjointStateInversion.run with _line_search and
FaultRecoveryPolicy (src/python/seismic_tomography/inversion.py:20-88,
124-272, 435-462) on top of the travel-time oracle, without the HDF5 writer.
Smoothing is SparseOps.smooth_gradient semantics (oracle/sparse.py)."""
from __future__ import annotations

import math

import numpy as np

from . import sparse as osp
from . import traveltime as ott


def tikhonov(velocity, gradient, nx, ny, nz, lam):
    v = np.asarray(velocity, np.float64).reshape(nx, ny, nz)
    lap = np.zeros_like(v)
    lap[1:-1, :, :] += v[2:, :, :] - 2 * v[1:-1, :, :] + v[:-2, :, :]
    lap[:, 1:-1, :] += v[:, 2:, :] - 2 * v[:, 1:-1, :] + v[:, :-2, :]
    lap[:, :, 1:-1] += v[:, :, 2:] - 2 * v[:, :, 1:-1] + v[:, :, :-2]
    return np.asarray(gradient, np.float64).reshape(-1) - lam * lap.ravel()


def normalize(gradient):
    g = np.asarray(gradient, np.float64).reshape(-1)
    m = np.max(np.abs(g))
    return g / m if m > 1e-30 else g.copy()


def depth_weights(nx, ny, nz, water_depth_iz, dz=200.0):
    p = np.ones((nx, ny, nz))
    for iz in range(nz):
        p[:, :, iz] = 0.0 if iz < water_depth_iz else 1.0 / (1.0 + (iz - water_depth_iz) * dz / 10000.0)
    return p.ravel()


def clamp_velocity(vp, lo=1450.0, hi=8500.0):
    return np.clip(np.where(np.isfinite(vp), vp, lo), lo, hi)


def invert(src, sta, n, origin, h, weights, observed, vp0, max_iterations=50, step_length=0.01, sigma=2.0,
           lam=1e-4, tol=1e-6, line_search_max=10, line_search_factor=0.5, descent=False):
    """The loop with fault tolerance on and a healthy model (no repairs):
    returns (vp, misfit history, converged)."""
    nx, ny, nz = n
    w = np.asarray(weights, np.float64)
    obs = np.asarray(observed, np.float64)
    vp = np.where(np.isfinite(vp0) & (vp0 > 0), vp0, 1500.0).astype(np.float64).ravel()

    def misfit(v):
        t = ott.forward(src, sta, n, origin, h, v)
        return 0.5 * float(np.sum(w * (obs - t) ** 2)), t
    hist, converged = [], False
    for it in range(max_iterations):
        J, t = misfit(vp)
        hist.append(J)
        if it > 0 and abs(hist[-2] - J) / max(abs(hist[-2]), 1e-30) < tol:
            converged = True
            break
        g = ott.gradient(src, sta, n, origin, h, vp, (obs - t) * w, w)
        g = np.where(np.isfinite(g), g, 0.0)
        g = osp.smooth_gradient(g, nx, ny, nz, sigma)
        g = tikhonov(vp, g, nx, ny, nz, lam)
        g = normalize(g)
        if descent:
            g = -g
        step = step_length
        for _ in range(line_search_max):
            Jt, _ = misfit(clamp_velocity(vp - step * g))
            if not math.isfinite(Jt):
                step *= 0.5
                continue
            if Jt < J:
                break
            step *= line_search_factor
        vp = clamp_velocity(vp - step * g)
    return vp, hist, converged
