// Gradient post-processing of the reference's GradientComputer
// (src/python/seismic_tomography/gradient.py:77-139), FP64 on the device.
//
//   regularization  g <- g - lambda * L(v), L the sum over axes of the second
//                   difference v[i+1] - 2 v[i] + v[i-1], taken only where both
//                   neighbours along that axis exist (an axis contributes 0 at
//                   its first and last index), :112-126
//   normalization   g <- g / max|g| when max|g| > 1e-30, :128-139
//   depth weights   p[., ., iz] = 0 above water_depth_iz, else
//                   1 / (1 + (iz - water_depth_iz) * dz / 10000), :96-110
//   preconditioning g <- g * p, :79-94
// Arrays are x-major (index (ix*ny + iy)*nz + iz).  The max reduction is
// exact (order-independent), so every result equals the reference's.  The
// volume kernels take z from the thread and y, x from the block grid
// (launch3.h); the max is reduced in two stages without atomics.
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

// A thread walks the rows of its block's chunk with the previous, current
// and next values of its z column in registers (the y neighbours) and takes
// its z neighbours from the neighbouring lanes; only the first and last lane
// of a wavefront load theirs.  Per cell: the next row's value, the two x
// neighbours and g.  Lanes past nz take part in the exchanges and store
// nothing.
__global__ void __launch_bounds__(256) tikhonov_kernel(int64_t nx, int64_t ny, int64_t nz, double lam,
                                                        const double* __restrict__ v, double* __restrict__ g) {
    const int64_t iz = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = iz < nz;
    const int64_t zc = act ? iz : nz - 1;   // a valid column for the lanes past nz
    const bool zin = act && iz > 0 && iz < nz - 1;
    const int lane = threadIdx.x & 63;
    const int64_t plane = ny * nz;
    for (int64_t ix = blockIdx.z; ix < nx; ix += gridDim.z) {
        const bool xin = ix > 0 && ix < nx - 1;
        for (int64_t y0 = (int64_t)blockIdx.y * stp::ROWS; y0 < ny; y0 += (int64_t)gridDim.y * stp::ROWS) {
            const int64_t y1 = y0 + stp::ROWS < ny ? y0 + stp::ROWS : ny;
            const int64_t base = ix * plane + zc;
            double vm = y0 > 0 ? v[base + (y0 - 1) * nz] : 0.0;
            double vc = v[base + y0 * nz];
            for (int64_t iy = y0; iy < y1; ++iy) {
                const int64_t i = base + iy * nz;
                const double vp = iy + 1 < ny ? v[i + nz] : 0.0;
                double zm = __shfl_up(vc, 1), zp = __shfl_down(vc, 1);
                if (lane == 0 && zin) zm = v[i - 1];
                if ((lane == 63 || iz + 1 >= nz) && zin) zp = v[i + 1];
                const double xm = xin ? v[i - plane] : 0.0, xp = xin ? v[i + plane] : 0.0;
                double lap = 0.0;
                // same association as the reference: per axis (v+ - 2v) + v-, axes added x, y, z
                if (xin) lap += xp - 2 * vc + xm;
                if (iy > 0 && iy < ny - 1) lap += vp - 2 * vc + vm;
                if (zin) lap += zp - 2 * vc + zm;
                if (act) g[i] = g[i] - lam * lap;
                vm = vc;
                vc = vp;
            }
        }
    }
}

// NaN propagates, as numpy's max does
__device__ __forceinline__ double max_nan(double a, double b) { return (b > a || b != b) ? b : a; }

// one partial max |g| per block; a final one-block pass reduces the partials
__device__ __forceinline__ double block_max(double m) {
    __shared__ double red[4];
    for (int o = 32; o > 0; o >>= 1) m = max_nan(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    return max_nan(max_nan(red[0], red[1]), max_nan(red[2], red[3]));
}

__global__ void __launch_bounds__(256) maxabs_kernel(int64_t n, const double* __restrict__ g, double* part) {
    double m = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        m = max_nan(m, fabs(g[i]));
    m = block_max(m);
    if (threadIdx.x == 0) part[blockIdx.x] = m;
}

__global__ void __launch_bounds__(256) max_final_kernel(int np, const double* part, double* mx) {
    double m = 0.0;
    for (int i = threadIdx.x; i < np; i += 256) m = max_nan(m, part[i]);
    m = block_max(m);
    if (threadIdx.x == 0) *mx = m;
}

__global__ void scale_by_max_kernel(int64_t n, const double* mx, double* g) {
    const double m = *mx;
    if (!(m > 1e-30)) return;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        g[i] = g[i] / m;
}

__global__ void depth_weights_kernel(int64_t nx, int64_t ny, int64_t nz, int64_t wd, double dz, double* p) {
    const int64_t iz = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (iz >= nz) return;
    const double w = iz < wd ? 0.0 : 1.0 / (1.0 + (double)(iz - wd) * dz / 10000.0);
    STP_FOR_PLANES_ROWS(ix, iy, nx, ny, stp::ROWS) p[(ix * ny + iy) * nz + iz] = w;
}

__global__ void multiply_kernel(int64_t n, const double* __restrict__ p, double* __restrict__ g) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        g[i] = g[i] * p[i];
}

bool dims_ok(int64_t nx, int64_t ny, int64_t nz) {
    return nx >= 1 && ny >= 1 && nz >= 1 && nx * ny <= ((int64_t)1 << 62) / nz;
}

}  // namespace

extern "C" {

int stp_gradient_tikhonov_f64(int64_t nx, int64_t ny, int64_t nz, double lambda, const double* velocity,
                              double* gradient, void* stream) {
    if (!dims_ok(nx, ny, nz) || !velocity || !gradient)
        return stp::fail(STP_ERR_INVALID, "stp_gradient_tikhonov_f64: bad argument");
    hipLaunchKernelGGL(tikhonov_kernel, stp::grid3(nx, ny, nz), dim3(256), 0, stp::S(stream), nx, ny, nz, lambda,
                       velocity, gradient);
    STP_HIP(hipGetLastError());
    return STP_OK;
}

int stp_gradient_normalize_f64(int64_t n, double* gradient, double* max_abs_host, void* stream) {
    if (n < 0 || (n > 0 && !gradient)) return stp::fail(STP_ERR_INVALID, "stp_gradient_normalize_f64: bad argument");
    if (max_abs_host) *max_abs_host = 0.0;
    if (n == 0) return STP_OK;
    hipStream_t st = stp::S(stream);
    const int64_t nb64 = (n + 255) / 256;
    const int np = (int)(nb64 < 2048 ? nb64 : 2048);
    double* buf = nullptr;   // np partials, then the max
    STP_HIP(hipMallocAsync(reinterpret_cast<void**>(&buf), sizeof(double) * (size_t)(np + 1), st));
    hipLaunchKernelGGL(maxabs_kernel, dim3(np), dim3(256), 0, st, n, gradient, buf);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        hipLaunchKernelGGL(max_final_kernel, dim3(1), dim3(256), 0, st, np, buf, buf + np);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(scale_by_max_kernel, dim3(blocks_for(n)), dim3(256), 0, st, n, buf + np, gradient);
        e = hipGetLastError();
    }
    double m = 0.0;
    if (e == hipSuccess && max_abs_host) e = hipMemcpyAsync(&m, buf + np, sizeof m, hipMemcpyDeviceToHost, st);
    hipError_t e2 = hipFreeAsync(buf, st);
    if (e == hipSuccess && max_abs_host) e = hipStreamSynchronize(st);
    if (e == hipSuccess) e = e2;
    if (e != hipSuccess) return stp::hip_fail(e, "stp_gradient_normalize_f64");
    if (max_abs_host) *max_abs_host = m;
    return STP_OK;
}

int stp_depth_preconditioner_f64(int64_t nx, int64_t ny, int64_t nz, int64_t water_depth_iz, double dz,
                                 double* out, void* stream) {
    if (!dims_ok(nx, ny, nz) || !out) return stp::fail(STP_ERR_INVALID, "stp_depth_preconditioner_f64: bad argument");
    hipLaunchKernelGGL(depth_weights_kernel, stp::grid3(nx, ny, nz), dim3(256), 0, stp::S(stream), nx, ny, nz,
                       water_depth_iz, dz, out);
    STP_HIP(hipGetLastError());
    return STP_OK;
}

int stp_gradient_precondition_f64(int64_t n, const double* weights, double* gradient, void* stream) {
    if (n < 0 || (n > 0 && (!weights || !gradient)))
        return stp::fail(STP_ERR_INVALID, "stp_gradient_precondition_f64: bad argument");
    if (n == 0) return STP_OK;
    hipLaunchKernelGGL(multiply_kernel, dim3(blocks_for(n)), dim3(256), 0, stp::S(stream), n, weights, gradient);
    STP_HIP(hipGetLastError());
    return STP_OK;
}

}  // extern "C"
