"""Synthetic inputs for the propagator, imaging and tomography workloads.

Everything here is input generation (host side or torch on the device), not
the measured path.  The white noise comes from a counter-based generator
(splitmix64 of seed and flat index) so any subset of the grid can be
regenerated independently and CPU and GPU runs see the same field.
"""
from __future__ import annotations

import math

import numpy as np

_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_G = np.uint64(0x9E3779B97F4A7C15)


def _splitmix64(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = x + _G
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, start: int, n: int, stream: int = 0) -> np.ndarray:
    """U(0,1) doubles for counters start..start+n-1 (53-bit, never 0)."""
    base = np.uint64((seed * 0x100000001B3 + stream * 0x9E3779B1) & 0xFFFFFFFFFFFFFFFF)
    idx = np.arange(start, start + n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        bits = _splitmix64(_splitmix64(idx ^ base) + base)
    return ((bits >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def normal(seed: int, n: int, stream: int = 0) -> np.ndarray:
    """N(0,1) by Box-Muller on two counter streams."""
    u1 = uniform(seed, 0, n, stream * 2 + 1)
    u2 = uniform(seed, 0, n, stream * 2 + 2)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


def _s64(c: int) -> int:
    c &= 0xFFFFFFFFFFFFFFFF
    return c - (1 << 64) if c >= (1 << 63) else c


def _torch_splitmix64(z):
    import torch
    def srl(x, k):  # logical right shift on int64 bit patterns
        return torch.bitwise_and(torch.bitwise_right_shift(x, k), (1 << (64 - k)) - 1)
    z = z + _s64(0x9E3779B97F4A7C15)
    z = torch.bitwise_xor(z, srl(z, 30)) * _s64(0xBF58476D1CE4E5B9)
    z = torch.bitwise_xor(z, srl(z, 27)) * _s64(0x94D049BB133111EB)
    return torch.bitwise_xor(z, srl(z, 31))


def uniform_torch(seed: int, n: int, stream: int = 0, device="cpu"):
    """Same values as uniform(seed, 0, n, stream), computed with torch (float64)."""
    import torch
    base = _s64(seed * 0x100000001B3 + stream * 0x9E3779B1)
    idx = torch.arange(n, dtype=torch.int64, device=device)
    bits = _torch_splitmix64(_torch_splitmix64(torch.bitwise_xor(idx, base)) + base)
    hi = torch.bitwise_and(torch.bitwise_right_shift(bits, 11), (1 << 53) - 1)
    return (hi.double() + 0.5) * (1.0 / 9007199254740992.0)


def normal_torch(seed: int, n: int, stream: int = 0, device="cpu"):
    import torch
    u1 = uniform_torch(seed, n, stream * 2 + 1, device)
    u2 = uniform_torch(seed, n, stream * 2 + 2, device)
    return torch.sqrt(-2.0 * torch.log(u1)) * torch.cos(2.0 * math.pi * u2)


def ricker(f: float, dt: float, n: int, t0: float | None = None) -> np.ndarray:
    """Ricker wavelet (1 - 2 pi^2 f^2 tau^2) exp(-pi^2 f^2 tau^2), tau = t - t0,
    t0 = 1/f by default (peak of value 1 at t0)."""
    if t0 is None:
        t0 = 1.0 / f
    tau = np.arange(n) * dt - t0
    a = (math.pi * f * tau) ** 2
    return (1.0 - 2.0 * a) * np.exp(-a)


def two_layer_velocity(nx: int, ny: int, nz: int, v_top: float = 1500.0, v_bottom: float = 2500.0,
                       split_iz: int | None = None) -> np.ndarray:
    """Interior velocity (nx, ny, nz): v_top for iz < split_iz, else v_bottom
    (depth along z; split at nz/2 by default, i.e. iz = 64 at 128^3)."""
    if split_iz is None:
        split_iz = nz // 2
    v = np.full((nx, ny, nz), v_bottom, dtype=np.float64)
    v[:, :, :split_iz] = v_top
    return v


def gaussian_kernel1d(sigma: float, truncate: float = 4.0) -> np.ndarray:
    r = int(truncate * sigma + 0.5)
    x = np.arange(-r, r + 1, dtype=np.float64)
    k = np.exp(-0.5 * (x / sigma) ** 2)
    return k / k.sum()


def smoothing_matrix(n: int, sigma: float, truncate: float = 4.0) -> np.ndarray:
    """(n, n) matrix of 1-D Gaussian smoothing with mirror boundaries
    (index -k maps to k, n-1+k to n-1-k, repeated for wide kernels)."""
    k = gaussian_kernel1d(sigma, truncate)
    r = (k.size - 1) // 2
    m = np.zeros((n, n))
    period = 2 * (n - 1) if n > 1 else 1
    for i in range(n):
        j = np.arange(i - r, i + r + 1)
        if n > 1:
            j = np.mod(j, period)
            j = np.where(j >= n, period - j, j)
        else:
            j = np.zeros_like(j)
        np.add.at(m[i], j, k)
    return m


def smooth_random_velocity(nx: int, ny: int, nz: int, seed: int = 2026, sigma: float = 20.0,
                           vmin: float = 1500.0, vmax: float = 4500.0, device: str = "cpu",
                           as_tensor: bool = False):
    """Seeded white noise (counter-based, float64) smoothed by a separable
    Gaussian of standard deviation `sigma` cells with mirror boundaries, then
    min-max scaled to [vmin, vmax].  The smoothing along each axis is a dense
    (n x n) matrix product.  Returns float64 (nx, ny, nz): numpy, or a torch
    tensor on `device` when as_tensor."""
    import torch
    t = normal_torch(seed, nx * ny * nz, device=device).reshape(nx, ny, nz)
    for axis, n in enumerate((nx, ny, nz)):
        m = torch.from_numpy(smoothing_matrix(n, sigma)).to(device)
        t = torch.movedim(torch.movedim(t, axis, -1) @ m.T, -1, axis).contiguous()
    lo, hi = t.min(), t.max()
    t = vmin + (t - lo) * ((vmax - vmin) / (hi - lo))
    return t if as_tensor else t.cpu().numpy()


def pad_halo(v_int: np.ndarray, hw: int) -> np.ndarray:
    """With-halo array from an interior array by edge replication (the halo
    velocity only enters the CFL maximum)."""
    return np.pad(v_int, hw, mode="edge")


def surface_receivers(nx: int, ny: int, nz: int, hw: int, iz: int = 0) -> np.ndarray:
    """With-halo flat indices of every interior cell of the plane iz (C order
    over ix, iy): receivers at every surface cell."""
    nyh, nzh = ny + 2 * hw, nz + 2 * hw
    ix, iy = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    return (((ix + hw) * nyh + (iy + hw)) * nzh + (iz + hw)).reshape(-1).astype(np.int64)


def line_receivers(nx: int, ny: int, nz: int, hw: int, iy: int, iz: int = 0, step: int = 2) -> np.ndarray:
    """Receivers on the line (ix = 0, step, ...), fixed iy and iz."""
    nyh, nzh = ny + 2 * hw, nz + 2 * hw
    ix = np.arange(0, nx, step)
    return (((ix + hw) * nyh + (iy + hw)) * nzh + (iz + hw)).astype(np.int64)


def cell_index(n: tuple[int, int, int], hw: int, ix: int, iy: int, iz: int) -> int:
    """With-halo flat index of interior cell (ix, iy, iz)."""
    nx, ny, nz = n
    return ((ix + hw) * (ny + 2 * hw) + (iy + hw)) * (nz + 2 * hw) + (iz + hw)


def ray_endpoints(nrays: int, n: tuple[int, int, int], h: float = 1.0, seed: int = 2026) -> np.ndarray:
    """(nrays, 2, 3) float64 source/receiver points on opposite faces of the
    box [0, n*h]^3.  Ray r joins face pair r % 3 (x, y, z faces); source on
    the low face, receiver on the high face, the other two coordinates of
    each point uniform from the counter-based generator."""
    ext = np.asarray(n, dtype=np.float64) * h
    u = np.stack([uniform(seed, 0, nrays, stream=10 + k) for k in range(4)], axis=1)
    pts = np.zeros((nrays, 2, 3))
    axis = np.arange(nrays) % 3
    for a in range(3):
        sel = axis == a
        o = [b for b in range(3) if b != a]
        pts[sel, 0, a] = 0.0
        pts[sel, 1, a] = ext[a]
        pts[sel, 0, o[0]] = u[sel, 0] * ext[o[0]]
        pts[sel, 0, o[1]] = u[sel, 1] * ext[o[1]]
        pts[sel, 1, o[0]] = u[sel, 2] * ext[o[0]]
        pts[sel, 1, o[1]] = u[sel, 3] * ext[o[1]]
    return pts
