"""Travel-time tomography on the GPU: the straight-ray forward model and
Frechet gradient of the reference's Python inversion
(src/python/seismic_tomography/inversion.py:329-433), FP64, through
stp_tt_forward / stp_tt_gradient of libseismic_hip.so.

The reference walks `obs_data.picks`, looks up each pick's source and station,
and loops over the samples of the segment between them.  Here the picks are
flattened once into device arrays of source and station coordinates (picks
whose source or station is missing keep a travel time of 0 and no gradient
contribution, as in the reference), and one GPU thread handles one pick.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import call, ptr, stream_handle


def _torch():
    import torch
    return torch


@dataclass
class Geometry:
    """Cell grid of the velocity model (inversion.py uses model.nx/ny/nz,
    dx/dy/dz and origin_x/y/z): vp has shape (nx, ny, nz), x-major."""
    nx: int
    ny: int
    nz: int
    dx: float = 1.0
    dy: float = 1.0
    dz: float = 1.0
    origin_x: float = 0.0
    origin_y: float = 0.0
    origin_z: float = 0.0

    def arrays(self):
        return (np.array([self.nx, self.ny, self.nz], np.int64),
                np.array([self.origin_x, self.origin_y, self.origin_z], np.float64),
                np.array([self.dx, self.dy, self.dz], np.float64))


class TravelTimeOperator:
    """Picks (source and station coordinates per pick, optional weights) on a
    model geometry.  `forward(vp)` is AdjointStateInversion._forward_model,
    `gradient(vp, weighted_residuals)` is _compute_gradient; misfit and
    residuals follow OBSDataCoordinator (obs_data.py:99-128)."""

    def __init__(self, geometry: Geometry, src_xyz, sta_xyz, weights=None, valid=None, device="cuda"):
        torch = _torch()
        self.geometry = geometry
        src = np.ascontiguousarray(np.asarray(src_xyz, np.float64).reshape(-1, 3))
        sta = np.ascontiguousarray(np.asarray(sta_xyz, np.float64).reshape(-1, 3))
        if src.shape != sta.shape:
            raise ValueError("src_xyz and sta_xyz must have the same number of picks")
        self.npicks = src.shape[0]
        self.valid = np.ones(self.npicks, bool) if valid is None else np.asarray(valid, bool).reshape(-1)
        keep = np.nonzero(self.valid)[0]
        self._keep = torch.as_tensor(keep, device=device)
        self.device = torch.device(device)
        self.src = torch.as_tensor(src[keep], device=device)
        self.sta = torch.as_tensor(sta[keep], device=device)
        w = np.ones(self.npicks) if weights is None else np.asarray(weights, np.float64).reshape(-1)
        self.weights = torch.as_tensor(w, device=device)
        self._n, self._o, self._h = geometry.arrays()

    @classmethod
    def from_obs(cls, geometry: Geometry, obs, device="cuda"):
        """From an object with the reference's OBSDataCoordinator shape:
        `picks` (with source_id, station_id, weight), `sources` and
        `stations` (dicts of objects with x, y, z)."""
        src = np.zeros((len(obs.picks), 3))
        sta = np.zeros((len(obs.picks), 3))
        valid = np.zeros(len(obs.picks), bool)
        for i, p in enumerate(obs.picks):
            s = obs.sources.get(p.source_id)
            r = obs.stations.get(p.station_id)
            if s is None or r is None:
                continue
            src[i] = (s.x, s.y, s.z)
            sta[i] = (r.x, r.y, r.z)
            valid[i] = True
        return cls(geometry, src, sta, weights=[p.weight for p in obs.picks], valid=valid, device=device)

    def _vp(self, vp):
        torch = _torch()
        g = self.geometry
        v = torch.as_tensor(vp, dtype=torch.float64, device=self.device).reshape(-1).contiguous()
        if v.numel() != g.nx * g.ny * g.nz:
            raise ValueError(f"vp has {v.numel()} cells, the geometry {g.nx * g.ny * g.nz}")
        return v

    def forward(self, vp):
        """Travel time per pick (FP64 device tensor, 0 for invalid picks)."""
        torch = _torch()
        v = self._vp(vp)
        nk = self.src.shape[0]
        tk = torch.empty(nk, dtype=torch.float64, device=self.device)
        call("stp_tt_forward", C.c_int64(nk), ptr(self.src), ptr(self.sta), self._n.ctypes.data, self._o.ctypes.data,
             self._h.ctypes.data, ptr(v), ptr(tk), stream_handle(self.device))
        t = torch.zeros(self.npicks, dtype=torch.float64, device=self.device)
        t[self._keep] = tk
        return t

    def gradient(self, vp, weighted_residuals):
        """Gradient per cell (FP64, flat x-major) for the weighted residuals
        the reference passes to _compute_gradient."""
        torch = _torch()
        v = self._vp(vp)
        r = torch.as_tensor(weighted_residuals, dtype=torch.float64, device=self.device).reshape(-1)
        rk = r[self._keep].contiguous()
        wk = self.weights[self._keep].contiguous()
        g = torch.empty(v.numel(), dtype=torch.float64, device=self.device)
        call("stp_tt_gradient", C.c_int64(rk.numel()), ptr(self.src), ptr(self.sta), self._n.ctypes.data,
             self._o.ctypes.data, self._h.ctypes.data, ptr(v), ptr(rk), ptr(wk), ptr(g), stream_handle(self.device))
        return g

    # ---------------------------------------------------------------- OBSDataCoordinator arithmetic
    def residuals(self, times, observed):
        """observed - calculated (compute_travel_time_residuals)."""
        torch = _torch()
        return torch.as_tensor(observed, dtype=torch.float64, device=self.device) - times

    def weighted_residuals(self, times, observed):
        return self.residuals(times, observed) * self.weights

    def misfit(self, times, observed, misfit_type: str = "l2") -> float:
        r = self.residuals(times, observed)
        if misfit_type in ("l2", "cross_correlation"):
            return float(0.5 * (self.weights * r * r).sum())
        if misfit_type == "l1":
            return float((self.weights * r.abs()).sum())
        raise ValueError(f"Unknown misfit type: {misfit_type}")
