"""TerrainBoundary of the reference's solver (src/cpp/solver/boundary.h:15-52,
boundary.cpp) on the GPU: the bathymetry grid lives on the device and the
fields are FP64 with-halo x-major arrays (device tensors are used in place,
numpy arrays are copied in and the result copied back, like the other
drop-in calls)."""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from ._lib import call, ptr, stream_handle


def _torch():
    import torch
    return torch


@dataclass
class BathymetryPoint:          # boundary.h:9-13
    x: float
    y: float
    depth: float


class TerrainBoundary:
    """`grid`: an object with the reference's GridParams fields (nx_global,
    ny_global, nz_global, dx, dy, dz, origin_x/y/z, halo_width); on one GPU
    the local block is the whole grid (offsets 0)."""

    def __init__(self, grid, water_depth: float):
        torch = _torch()
        self.grid = grid
        self.water_depth = float(water_depth)
        self.nx, self.ny, self.nz = int(grid.nx_global), int(grid.ny_global), int(grid.nz_global)
        self.hw = int(grid.halo_width)
        self.bathy = torch.full((self.nx * self.ny,), self.water_depth, dtype=torch.float64, device="cuda")
        self._mask = None

    # ------------------------------------------------------------ geometry
    def _ext(self):
        h = self.hw
        return np.array([self.nx + 2 * h, self.ny + 2 * h, self.nz + 2 * h, h, 0, 0, 0, self.nx, self.ny], np.int64)

    def _geo(self):
        return np.array([self.water_depth, float(self.grid.dz), float(self.grid.origin_z)], np.float64)

    def bathymetry(self):
        return self.bathy

    def set_bathymetry_data(self, data) -> None:
        """boundary.cpp:15-31: each point sets the cell it falls in (later
        points overwrite earlier ones); points outside the grid are ignored."""
        g = self.grid
        host = self.bathy.cpu().numpy()
        for pt in data:
            ix = int((pt.x - g.origin_x) / g.dx)
            iy = int((pt.y - g.origin_y) / g.dy)
            if 0 <= ix < self.nx and 0 <= iy < self.ny:
                host[ix * self.ny + iy] = pt.depth
        self.bathy.copy_(_torch().as_tensor(host))

    def generate_synthetic_bathymetry(self, gradient: float, roughness: float) -> None:
        call("stp_terrain_bathymetry_synthetic_f64", C.c_int64(self.nx), C.c_int64(self.ny),
             C.c_double(self.grid.dx), C.c_double(self.grid.dy), C.c_double(self.water_depth),
             C.c_double(gradient), C.c_double(roughness), ptr(self.bathy), stream_handle())

    def seafloor_depth(self, gx: int, gy: int) -> float:
        if gx < 0 or gx >= self.nx or gy < 0 or gy >= self.ny:
            return self.water_depth
        return float(self.bathy[gx * self.ny + gy])

    def seafloor_index(self, gx: int, gy: int) -> int:
        return int(self.seafloor_depth(gx, gy) / self.grid.dz)

    def is_subsurface(self, gx: int, gy: int, gz: int) -> bool:
        return self.grid.origin_z + gz * self.grid.dz >= self.seafloor_depth(gx, gy)

    def is_water(self, gx: int, gy: int, gz: int) -> bool:
        return not self.is_subsurface(gx, gy, gz)

    # ------------------------------------------------------------ field operations
    @staticmethod
    def _field(field):
        torch = _torch()
        if torch.is_tensor(field):
            if field.dtype != torch.float64 or not field.is_cuda or not field.is_contiguous():
                raise TypeError("tensor fields must be contiguous float64 CUDA tensors (they are modified in place)")
            return field.reshape(-1), None
        host = np.ascontiguousarray(np.asarray(field, np.float64)).reshape(-1)
        return torch.as_tensor(host).to("cuda"), host

    @staticmethod
    def _back(field, dev, host):
        if host is None:
            return field
        out = dev.cpu().numpy()
        if isinstance(field, np.ndarray) and field.dtype == np.float64 and field.flags.c_contiguous:
            field.reshape(-1)[:] = out
            return field
        return out

    def _check(self, d):
        h = self.hw
        n = (self.nx + 2 * h) * (self.ny + 2 * h) * (self.nz + 2 * h)
        if d.numel() < n:
            raise ValueError(f"field has {d.numel()} values, the grid with halo {n}")

    def update_velocity_mask(self, velocity, water_vp: float, sediment_vp: float):
        d, host = self._field(velocity)
        self._check(d)
        call("stp_terrain_velocity_mask_f64", self._ext().ctypes.data, self._geo().ctypes.data, ptr(self.bathy),
             C.c_double(water_vp), C.c_double(sediment_vp), ptr(d), stream_handle())
        return self._back(velocity, d, host)

    def build_boundary_mask(self, loc=None, halo_width=None):
        torch = _torch()
        h = self.hw
        m = torch.empty((self.nx + 2 * h) * (self.ny + 2 * h) * (self.nz + 2 * h), dtype=torch.uint8, device="cuda")
        call("stp_terrain_boundary_mask", self._ext().ctypes.data, self._geo().ctypes.data, ptr(self.bathy), ptr(m),
             stream_handle())
        self._mask = m
        return m

    def mask(self):
        return self._mask

    def apply_free_surface_bc(self, field, loc=None, halo_width=None):
        d, host = self._field(field)
        self._check(d)
        h = self.hw
        call("stp_terrain_free_surface_f64", C.c_int64(self.nx + 2 * h), C.c_int64(self.ny + 2 * h),
             C.c_int64(self.nz + 2 * h), C.c_int64(h), ptr(d), stream_handle())
        return self._back(field, d, host)

    def apply_seafloor_bc(self, field, loc=None, halo_width=None):
        d, host = self._field(field)
        self._check(d)
        call("stp_terrain_seafloor_f64", self._ext().ctypes.data, self._geo().ctypes.data, ptr(self.bathy), ptr(d),
             C.c_int64(d.numel()), stream_handle())
        return self._back(field, d, host)

    def apply_absorbing_bc(self, field, loc=None, halo_width=None, sponge_width: int = 20, damping: float = 0.015):
        d, host = self._field(field)
        self._check(d)
        h = self.hw
        call("stp_terrain_absorbing_f64", C.c_int64(self.nx + 2 * h), C.c_int64(self.ny + 2 * h),
             C.c_int64(self.nz + 2 * h), C.c_int64(h), int(sponge_width), C.c_double(damping), ptr(d),
             stream_handle())
        return self._back(field, d, host)
