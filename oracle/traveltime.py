"""CPU oracle of the straight-ray travel-time operator.  TEST INFRASTRUCTURE
ONLY: imported by tests/ and never by the product path.

Restates AdjointStateInversion._forward_model and _compute_gradient
(src/python/seismic_tomography/inversion.py:329-383 and :385-433) with the
reference's scalar FP64 arithmetic (Python floats, same operation order), one
pick at a time:
  * cell of a point: int((coord - origin) / h) truncated toward zero, clamped
    to [0, n - 1] (inversion.py:340-346);
  * samples: n_s = max(int(dist / (dx * 5)), 2), frac = (s + 0.5) / n_s, cell
    int(i_src + frac * (i_sta - i_src)) clamped (:348-359), ds = dist / n_s;
  * forward: travel time += ds / v, v non-finite or <= 0 replaced by 1500, a
    non-finite contribution skipped, a non-finite or negative total replaced
    by dist / 6000 (:361-381);
  * gradient: += r_i * (-ds / (v * v)) * w_i, invalid v and non-finite
    contributions skipped, picks with a non-finite residual skipped (:404-431).
"""
from __future__ import annotations

import math

import numpy as np


def _cell(c, o, h, n):
    return max(0, min(int((c - o) / h), n - 1))


def _samples(src, sta, n, origin, h):
    d = math.sqrt((src[0] - sta[0]) ** 2 + (src[1] - sta[1]) ** 2 + (src[2] - sta[2]) ** 2)
    a = [_cell(src[k], origin[k], h[k], n[k]) for k in range(3)]
    b = [_cell(sta[k], origin[k], h[k], n[k]) for k in range(3)]
    ns = max(int(d / (h[0] * 5)), 2)
    ds = d / ns
    cells = []
    for s in range(ns):
        frac = (s + 0.5) / ns
        c = [max(0, min(int(a[k] + frac * (b[k] - a[k])), n[k] - 1)) for k in range(3)]
        cells.append((c[0] * n[1] + c[1]) * n[2] + c[2])
    return d, ds, cells


def forward(src_xyz, sta_xyz, n, origin, h, vp):
    """Travel time per pick; vp flat x-major (nx*ny*nz)."""
    vp = np.asarray(vp, np.float64).reshape(-1)
    src_xyz = np.asarray(src_xyz, np.float64).reshape(-1, 3)
    sta_xyz = np.asarray(sta_xyz, np.float64).reshape(-1, 3)
    out = np.zeros(len(src_xyz))
    for i in range(len(src_xyz)):
        d, ds, cells = _samples(src_xyz[i].tolist(), sta_xyz[i].tolist(), n, origin, h)
        t = 0.0
        for c in cells:
            v = float(vp[c])
            if not math.isfinite(v) or v <= 0:
                v = 1500.0
            x = ds / v
            if not math.isfinite(x):
                continue
            t += x
        if not math.isfinite(t) or t < 0:
            t = d / 6000.0
        out[i] = t
    return out


def gradient(src_xyz, sta_xyz, n, origin, h, vp, weighted_residuals, weights=None):
    vp = np.asarray(vp, np.float64).reshape(-1)
    src_xyz = np.asarray(src_xyz, np.float64).reshape(-1, 3)
    sta_xyz = np.asarray(sta_xyz, np.float64).reshape(-1, 3)
    r = np.asarray(weighted_residuals, np.float64).reshape(-1)
    w = np.ones(len(r)) if weights is None else np.asarray(weights, np.float64).reshape(-1)
    g = np.zeros(vp.size)
    for i in range(len(src_xyz)):
        if not math.isfinite(r[i]):
            continue
        _, ds, cells = _samples(src_xyz[i].tolist(), sta_xyz[i].tolist(), n, origin, h)
        for c in cells:
            v = float(vp[c])
            if not math.isfinite(v) or v <= 0:
                continue
            f = -ds / (v * v)
            if not math.isfinite(f):
                continue
            g[c] += float(r[i]) * f * float(w[i])
    return g
