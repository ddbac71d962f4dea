// Terrain (bathymetry) boundary of the reference's solver
// (src/cpp/solver/boundary.cpp:33-204), FP64 on the device.
//
//   synthetic bathymetry  depth(ix, iy) = water + g x + r sin(2 pi x / 5000)
//                         cos(2 pi y / 7000) + 0.5 r sin(4 pi x / 3000),
//                         clamped to [500, 8000] m (x = ix dx, y = iy dy)
//   seafloor index        int(depth / dz); columns outside the global grid
//                         have the water depth
//   velocity mask         water above the seafloor index, below it the
//                         sediment velocity times 1 + 0.3 (iz - sf) dz / 5000,
//                         capped at 8500 m/s
//   boundary mask         1 below the seafloor (z >= depth), 2 above, 0 halo
//   free surface          zero the first interior plane (iz = hw)
//   seafloor              zero iz = hw .. hw + sf - 1 of each column, at flat
//                         indices below the field size (the reference's loop
//                         is bounded by the field size only, so a seafloor
//                         deeper than the grid zeroes cells of the following
//                         columns; every write is a zero, so the order does not
//                         matter)
//   absorbing             the reference's sponge factor from the distance to
//                         the nearest interior face
// Arrays are with-halo x-major like the wavefields; the bathymetry is
// [nx_global][ny_global].  The arithmetic follows the reference's operation
// order with contraction off, so results other than the synthetic bathymetry
// (which uses the device sin/cos) are bit-identical.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "capi_common.h"
#include "launch3.h"
#include "seismic_hip.h"

namespace {

#pragma clang fp contract(off)

unsigned blocks_for(int64_t n) {
    const int64_t b = (n + 255) / 256;
    return (unsigned)(b < 1 ? 1 : (b > 65535 ? 65535 : b));
}

struct Column {
    int64_t nxh, nyh, nzh, hw, ox, oy, oz, nxg, nyg;
    double water_depth, dz, origin_z;
};

__device__ __forceinline__ double depth_at(const Column& c, const double* bathy, int64_t gx, int64_t gy) {
    if (gx < 0 || gx >= c.nxg || gy < 0 || gy >= c.nyg) return c.water_depth;
    return bathy[gx * c.nyg + gy];
}

__global__ void bathy_kernel(int64_t nx, int64_t ny, double dx, double dy, double water, double g, double r,
                             double* bathy) {
    const int64_t n = nx * ny;
    const double pi = 3.14159265358979323846;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t ix = i / ny, iy = i % ny;
        const double x = ix * dx, y = iy * dy;
        double d = water + g * x + r * sin(2.0 * pi * x / 5000.0) * cos(2.0 * pi * y / 7000.0) +
                   r * 0.5 * sin(4.0 * pi * x / 3000.0);
        d = fmax(d, 500.0);
        d = fmin(d, 8000.0);
        bathy[i] = d;
    }
}

// interior cells; launched over the interior box (launch3.h)
__global__ void velocity_mask_kernel(Column c, const double* bathy, double water_vp, double sediment_vp, double* v) {
    const int64_t nxi = c.nxh - 2 * c.hw, nyi = c.nyh - 2 * c.hw, nzi = c.nzh - 2 * c.hw;
    const int64_t tz = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tz >= nzi) return;
    const int64_t iz = tz + c.hw;
    STP_FOR_PLANES_ROWS(tx, ty, nxi, nyi, stp::ROWS) {
            const int64_t ix = tx + c.hw, iy = ty + c.hw;
            const int64_t sf = (int64_t)(depth_at(c, bathy, tx + c.ox, ty + c.oy) / c.dz);
            const int64_t idx = (ix * c.nyh + iy) * c.nzh + iz;
            if (iz - c.hw < sf) {
                v[idx] = water_vp;
            } else {
                const double grad = 1.0 + 0.3 * ((iz - c.hw - sf) * c.dz / 5000.0);
                v[idx] = fmin(sediment_vp * grad, 8500.0);
            }
        }
}

// every cell of the with-halo box (launch3.h)
__global__ void boundary_mask_kernel(Column c, const double* bathy, uint8_t* mask) {
    const int64_t iz = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (iz >= c.nzh) return;
    const bool zin = iz >= c.hw && iz < c.nzh - c.hw;
    const double zpos = c.origin_z + (iz - c.hw + c.oz) * c.dz;
    STP_FOR_PLANES_ROWS(ix, iy, c.nxh, c.nyh, stp::ROWS) {
            uint8_t m = 0;
            if (zin && ix >= c.hw && ix < c.nxh - c.hw && iy >= c.hw && iy < c.nyh - c.hw)
                m = zpos >= depth_at(c, bathy, ix - c.hw + c.ox, iy - c.hw + c.oy) ? 1 : 2;
            mask[(ix * c.nyh + iy) * c.nzh + iz] = m;
        }
}

__global__ void free_surface_kernel(int64_t nxh, int64_t nyh, int64_t nzh, int64_t hw, double* f) {
    const int64_t nxi = nxh - 2 * hw, nyi = nyh - 2 * hw;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nxi * nyi;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t ix = t / nyi + hw, iy = t % nyi + hw;
        f[(ix * nyh + iy) * nzh + hw] = 0.0;
    }
}

__global__ void seafloor_kernel(Column c, const double* bathy, int64_t size, double* f) {
    const int64_t nxi = c.nxh - 2 * c.hw, nyi = c.nyh - 2 * c.hw;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nxi * nyi;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t ix = t / nyi + c.hw, iy = t % nyi + c.hw;
        const int64_t sf = (int64_t)(depth_at(c, bathy, ix - c.hw + c.ox, iy - c.hw + c.oy) / c.dz);
        const int64_t base = (ix * c.nyh + iy) * c.nzh;
        for (int64_t iz = c.hw; iz < c.hw + sf; ++iz) {
            const int64_t idx = base + iz;
            if (idx >= size) break;
            f[idx] = 0.0;
        }
    }
}

// interior cells; launched over the interior box (launch3.h)
__global__ void absorbing_kernel(int64_t nxh, int64_t nyh, int64_t nzh, int64_t hw, int sw, double damping,
                                 double* f) {
    const int64_t nxi = nxh - 2 * hw, nyi = nyh - 2 * hw, nzi = nzh - 2 * hw;
    const int64_t tz = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tz >= nzi) return;
    const int64_t iz = tz + hw;
    STP_FOR_PLANES_ROWS(tx, ty, nxi, nyi, stp::ROWS) {
            const int64_t ix = tx + hw, iy = ty + hw;
            double d = (double)(iz - hw);
            d = fmin(d, (double)(nzh - hw - 1 - iz));
            d = fmin(d, (double)(ix - hw));
            d = fmin(d, (double)(nxh - hw - 1 - ix));
            d = fmin(d, (double)(iy - hw));
            d = fmin(d, (double)(nyh - hw - 1 - iy));
            if (d < sw && d >= 0) {
                const double factor = 1.0 - damping * (sw - d) * (sw - d) / (double)(sw * sw);
                const int64_t idx = (ix * nyh + iy) * nzh + iz;
                f[idx] *= factor;
            }
        }
}

bool column_of(const int64_t* ext, const double* geo, Column* c) {
    // ext: nxh, nyh, nzh, hw, offset x/y/z, nx_global, ny_global; geo: water depth, dz, origin_z
    if (!ext || !geo) return false;
    c->nxh = ext[0]; c->nyh = ext[1]; c->nzh = ext[2]; c->hw = ext[3];
    c->ox = ext[4]; c->oy = ext[5]; c->oz = ext[6]; c->nxg = ext[7]; c->nyg = ext[8];
    c->water_depth = geo[0]; c->dz = geo[1]; c->origin_z = geo[2];
    return c->hw >= 0 && c->nxh > 2 * c->hw && c->nyh > 2 * c->hw && c->nzh > 2 * c->hw && c->nxg >= 1 &&
           c->nyg >= 1 && c->dz > 0.0 && c->nxh * c->nyh <= ((int64_t)1 << 62) / c->nzh;
}

}  // namespace

extern "C" {

int stp_terrain_bathymetry_synthetic_f64(int64_t nx, int64_t ny, double dx, double dy, double water_depth,
                                         double gradient, double roughness, double* bathy, void* stream) {
    if (nx < 1 || ny < 1 || !bathy || nx > ((int64_t)1 << 62) / ny)
        return stp::fail(STP_ERR_INVALID, "stp_terrain_bathymetry_synthetic_f64: bad argument");
    hipLaunchKernelGGL(bathy_kernel, dim3(blocks_for(nx * ny)), dim3(256), 0, stp::S(stream), nx, ny, dx, dy,
                       water_depth, gradient, roughness, bathy);
    STP_HIP(hipGetLastError());
    return STP_OK;
}

int stp_terrain_velocity_mask_f64(const int64_t* ext, const double* geo, const double* bathy, double water_vp,
                                  double sediment_vp, double* velocity, void* stream) {
    Column c;
    if (!column_of(ext, geo, &c) || !bathy || !velocity)
        return stp::fail(STP_ERR_INVALID, "stp_terrain_velocity_mask_f64: bad argument");
    const dim3 g3 = stp::grid3(c.nxh - 2 * c.hw, c.nyh - 2 * c.hw, c.nzh - 2 * c.hw);
    hipLaunchKernelGGL(velocity_mask_kernel, g3, dim3(256), 0, stp::S(stream), c, bathy, water_vp, sediment_vp,
                       velocity);
    STP_HIP(hipGetLastError());
    return STP_OK;
}

int stp_terrain_boundary_mask(const int64_t* ext, const double* geo, const double* bathy, uint8_t* mask,
                              void* stream) {
    Column c;
    if (!column_of(ext, geo, &c) || !bathy || !mask)
        return stp::fail(STP_ERR_INVALID, "stp_terrain_boundary_mask: bad argument");
    hipLaunchKernelGGL(boundary_mask_kernel, stp::grid3(c.nxh, c.nyh, c.nzh), dim3(256), 0, stp::S(stream), c,
                       bathy, mask);
    STP_HIP(hipGetLastError());
    return STP_OK;
}

int stp_terrain_free_surface_f64(int64_t nxh, int64_t nyh, int64_t nzh, int64_t hw, double* field, void* stream) {
    if (hw < 0 || nxh <= 2 * hw || nyh <= 2 * hw || nzh <= 2 * hw || !field)
        return stp::fail(STP_ERR_INVALID, "stp_terrain_free_surface_f64: bad argument");
    hipLaunchKernelGGL(free_surface_kernel, dim3(blocks_for((nxh - 2 * hw) * (nyh - 2 * hw))), dim3(256), 0,
                       stp::S(stream), nxh, nyh, nzh, hw, field);
    STP_HIP(hipGetLastError());
    return STP_OK;
}

int stp_terrain_seafloor_f64(const int64_t* ext, const double* geo, const double* bathy, double* field,
                             int64_t field_size, void* stream) {
    Column c;
    if (!column_of(ext, geo, &c) || !bathy || !field || field_size < 0)
        return stp::fail(STP_ERR_INVALID, "stp_terrain_seafloor_f64: bad argument");
    hipLaunchKernelGGL(seafloor_kernel, dim3(blocks_for((c.nxh - 2 * c.hw) * (c.nyh - 2 * c.hw))), dim3(256), 0,
                       stp::S(stream), c, bathy, field_size, field);
    STP_HIP(hipGetLastError());
    return STP_OK;
}

int stp_terrain_absorbing_f64(int64_t nxh, int64_t nyh, int64_t nzh, int64_t hw, int32_t sponge_width,
                              double damping, double* field, void* stream) {
    if (hw < 0 || nxh <= 2 * hw || nyh <= 2 * hw || nzh <= 2 * hw || !field)
        return stp::fail(STP_ERR_INVALID, "stp_terrain_absorbing_f64: bad argument");
    hipLaunchKernelGGL(absorbing_kernel, stp::grid3(nxh - 2 * hw, nyh - 2 * hw, nzh - 2 * hw), dim3(256),
                       0, stp::S(stream), nxh, nyh, nzh, hw, (int)sponge_width, damping, field);
    STP_HIP(hipGetLastError());
    return STP_OK;
}

}  // extern "C"
