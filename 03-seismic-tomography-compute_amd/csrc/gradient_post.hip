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
// exact (order-independent), so every result equals the reference's.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "capi_common.h"
#include "seismic_hip.h"

namespace {

#pragma clang fp contract(off)

unsigned blocks_for(int64_t n) {
    const int64_t b = (n + 255) / 256;
    return (unsigned)(b < 1 ? 1 : (b > 65535 ? 65535 : b));
}

__global__ void tikhonov_kernel(int64_t nx, int64_t ny, int64_t nz, double lam, const double* __restrict__ v,
                                double* __restrict__ g) {
    const int64_t n = nx * ny * nz;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t iz = i % nz, iy = (i / nz) % ny, ix = i / (ny * nz);
        const double c = v[i];
        double lap = 0.0;
        // same association as the reference: per axis (v+ - 2v) + v-, axes added x, y, z
        if (ix > 0 && ix < nx - 1) lap += v[i + ny * nz] - 2 * c + v[i - ny * nz];
        if (iy > 0 && iy < ny - 1) lap += v[i + nz] - 2 * c + v[i - nz];
        if (iz > 0 && iz < nz - 1) lap += v[i + 1] - 2 * c + v[i - 1];
        g[i] = g[i] - lam * lap;
    }
}

__global__ void maxabs_kernel(int64_t n, const double* __restrict__ g, unsigned long long* out) {
    double m = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double a = fabs(g[i]);
        m = (a > m || a != a) ? a : m;   // NaN propagates, as numpy's max does
    }
    // non-negative doubles (and NaN, whose bit pattern with sign 0 is the
    // largest) order like their unsigned bit patterns
    for (int o = 32; o > 0; o >>= 1) {
        const double p = __shfl_xor(m, o);
        m = (p > m || p != p) ? p : m;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

__global__ void scale_by_max_kernel(int64_t n, const unsigned long long* mx, double* g) {
    const double m = __longlong_as_double((long long)*mx);
    if (!(m > 1e-30)) return;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        g[i] = g[i] / m;
}

__global__ void depth_weights_kernel(int64_t nx, int64_t ny, int64_t nz, int64_t wd, double dz, double* p) {
    const int64_t n = nx * ny * nz;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t iz = i % nz;
        p[i] = iz < wd ? 0.0 : 1.0 / (1.0 + (double)(iz - wd) * dz / 10000.0);
    }
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
    const int64_t n = nx * ny * nz;
    hipLaunchKernelGGL(tikhonov_kernel, dim3(blocks_for(n)), dim3(256), 0, stp::S(stream), nx, ny, nz, lambda,
                       velocity, gradient);
    STP_HIP(hipGetLastError());
    return STP_OK;
}

int stp_gradient_normalize_f64(int64_t n, double* gradient, double* max_abs_host, void* stream) {
    if (n < 0 || (n > 0 && !gradient)) return stp::fail(STP_ERR_INVALID, "stp_gradient_normalize_f64: bad argument");
    if (max_abs_host) *max_abs_host = 0.0;
    if (n == 0) return STP_OK;
    hipStream_t st = stp::S(stream);
    unsigned long long* mx = nullptr;
    STP_HIP(hipMallocAsync(reinterpret_cast<void**>(&mx), sizeof(unsigned long long), st));
    hipError_t e = hipMemsetAsync(mx, 0, sizeof(unsigned long long), st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(maxabs_kernel, dim3(blocks_for(n)), dim3(256), 0, st, n, gradient, mx);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(scale_by_max_kernel, dim3(blocks_for(n)), dim3(256), 0, st, n, mx, gradient);
        e = hipGetLastError();
    }
    unsigned long long bits = 0;
    if (e == hipSuccess && max_abs_host) e = hipMemcpyAsync(&bits, mx, sizeof bits, hipMemcpyDeviceToHost, st);
    hipError_t e2 = hipFreeAsync(mx, st);
    if (e == hipSuccess && max_abs_host) e = hipStreamSynchronize(st);
    if (e == hipSuccess) e = e2;
    if (e != hipSuccess) return stp::hip_fail(e, "stp_gradient_normalize_f64");
    if (max_abs_host) {
        double m;
        static_assert(sizeof m == sizeof bits, "size");
        __builtin_memcpy(&m, &bits, sizeof m);
        *max_abs_host = m;
    }
    return STP_OK;
}

int stp_depth_preconditioner_f64(int64_t nx, int64_t ny, int64_t nz, int64_t water_depth_iz, double dz,
                                 double* out, void* stream) {
    if (!dims_ok(nx, ny, nz) || !out) return stp::fail(STP_ERR_INVALID, "stp_depth_preconditioner_f64: bad argument");
    const int64_t n = nx * ny * nz;
    hipLaunchKernelGGL(depth_weights_kernel, dim3(blocks_for(n)), dim3(256), 0, stp::S(stream), nx, ny, nz,
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
