"""Terrain (bathymetry) boundary: the reference's TerrainBoundary compiled
from its own source (oracle/_ref, src/cpp/solver/boundary.cpp) against
closed forms on the CPU, and the GPU kernels (stp_terrain_*) against it:
masks, velocities and boundary conditions bit-exact, the synthetic
bathymetry (device sin/cos) within 1e-15 relative."""
import ctypes as C
from pathlib import Path
from types import SimpleNamespace

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
REFLIB = ROOT / "oracle" / "_ref" / "libref.so"
pytestmark = pytest.mark.skipif(not REFLIB.exists(), reason="reference build (oracle/_ref) not present")

N = (14, 11, 40)
HW = 2
H = (250.0, 300.0, 150.0)
O = (0.0, 0.0, 0.0)


def _d(a):
    return a.ctypes.data_as(C.c_void_p)


def ref_terrain(gen=True, pts=None, water=2000.0, field=None, sw=5, damping=0.3):
    lib = C.CDLL(str(REFLIB))
    nx, ny, nz = N
    nh = (nx + 2 * HW) * (ny + 2 * HW) * (nz + 2 * HW)
    pts = np.zeros((0, 3)) if pts is None else np.ascontiguousarray(pts, np.float64)
    field = np.ascontiguousarray(field, np.float64)
    out = dict(bathy=np.zeros(nx * ny), vel=np.zeros(nh), mask=np.zeros(nh, np.uint8), free=np.zeros(nh),
               sea=np.zeros(nh), absorb=np.zeros(nh))
    h = np.array(H)
    o = np.array(O)
    lib.ref_terrain(C.c_int64(nx), C.c_int64(ny), C.c_int64(nz), C.c_int(HW), _d(h), _d(o), C.c_double(water),
                    C.c_int(int(gen)), C.c_double(0.08), C.c_double(300.0), C.c_int64(len(pts)), _d(pts),
                    C.c_double(1500.0), C.c_double(1800.0), C.c_int(sw), C.c_double(damping), _d(field),
                    _d(out["bathy"]), _d(out["vel"]), _d(out["mask"]), _d(out["free"]), _d(out["sea"]),
                    _d(out["absorb"]))
    return out


def _field(seed=0):
    nh = (N[0] + 2 * HW) * (N[1] + 2 * HW) * (N[2] + 2 * HW)
    return np.random.default_rng(seed).standard_normal(nh)


def test_reference_bathymetry_and_velocity_closed_forms():
    out = ref_terrain(field=_field())
    nx, ny, nz = N
    x = np.arange(nx)[:, None] * H[0]
    y = np.arange(ny)[None, :] * H[1]
    d = 2000.0 + 0.08 * x + 300.0 * np.sin(2 * np.pi * x / 5000) * np.cos(2 * np.pi * y / 7000) + \
        150.0 * np.sin(4 * np.pi * x / 3000)
    np.testing.assert_allclose(out["bathy"], np.clip(d, 500, 8000).ravel(), rtol=1e-14)
    # velocity of column (3, 4): water above int(depth / dz), then 1800 (1 + 0.3 k dz / 5000)
    nyh, nzh = ny + 2 * HW, nz + 2 * HW
    sf = int(out["bathy"][3 * ny + 4] / H[2])
    col = out["vel"].reshape(nx + 2 * HW, nyh, nzh)[3 + HW, 4 + HW, HW:-HW]
    k = np.arange(nz) - sf
    expect = np.where(k < 0, 1500.0, np.minimum(1800.0 * (1 + 0.3 * (k * H[2] / 5000.0)), 8500.0))
    np.testing.assert_array_equal(col, expect)


@pytest.mark.gpu
@pytest.mark.parametrize("deep", [False, True])
def test_gpu_terrain_matches_reference(deep, cuda):
    import torch
    from seismic_engine.terrain import BathymetryPoint, TerrainBoundary
    field = _field(1)
    water = 9000.0 if deep else 2000.0          # deeper than the grid: the seafloor loop runs past the column
    pts = np.array([[260.0, 310.0, 123.0], [1000.0, 20.0, 4550.0], [-5.0, 10.0, 1.0], [1000.0, 20.0, 800.0]])
    ref = ref_terrain(gen=not deep, pts=pts, water=water, field=field)
    grid = SimpleNamespace(nx_global=N[0], ny_global=N[1], nz_global=N[2], dx=H[0], dy=H[1], dz=H[2],
                           origin_x=O[0], origin_y=O[1], origin_z=O[2], halo_width=HW)
    t = TerrainBoundary(grid, water)
    if not deep:
        t.generate_synthetic_bathymetry(0.08, 300.0)
    t.set_bathymetry_data([BathymetryPoint(*p) for p in pts])
    b = t.bathymetry().cpu().numpy()
    assert np.abs(b - ref["bathy"]).max() <= 1e-15 * np.abs(ref["bathy"]).max()
    # the remaining operations from the reference's bathymetry, so they are compared bit for bit
    t.bathy.copy_(torch.as_tensor(ref["bathy"]))
    assert np.array_equal(t.update_velocity_mask(field.copy(), 1500.0, 1800.0), ref["vel"])
    assert np.array_equal(t.build_boundary_mask().cpu().numpy(), ref["mask"])
    assert np.array_equal(t.apply_free_surface_bc(field.copy()), ref["free"])
    assert np.array_equal(t.apply_seafloor_bc(field.copy()), ref["sea"])
    assert np.array_equal(t.apply_absorbing_bc(field.copy(), sponge_width=5, damping=0.3), ref["absorb"])
    dev = torch.as_tensor(field, device=cuda)
    t.apply_absorbing_bc(dev, sponge_width=5, damping=0.3)
    assert np.array_equal(dev.cpu().numpy(), ref["absorb"])
    assert t.seafloor_index(3, 4) == int(ref["bathy"][3 * N[1] + 4] / H[2])
    assert t.is_water(3, 4, 0) == (0.0 < ref["bathy"][3 * N[1] + 4])
