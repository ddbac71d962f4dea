// RTM imaging primitives in FP64, the reference RTMEngine semantics
// (src/cpp/rtm/rtm_engine.cpp:24-57, 170-182) and the elementwise adjoint
// gradient of GradientComputer.compute_adjoint_gradient
// (src/python/seismic_tomography/gradient.py:24-40).  No FP contraction: each
// product and sum is rounded as in the reference's C++.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "capi_common.h"

#pragma clang fp contract(off)

namespace stp {
namespace {

__global__ void image_condition_kernel(int64_t nx, int64_t ny, int64_t nz, int hw, int64_t sy, int64_t sz,
                                       const double* f, const double* a, double* image, double* illum) {
    const int64_t n = nx * ny * nz;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t ix = i / (ny * nz), iy = (i / nz) % ny, iz = i % nz;
        const int64_t h = (ix + hw) * (sy * sz) + (iy + hw) * sz + (iz + hw);
        const double fv = __builtin_isfinite(f[h]) ? f[h] : 0.0;
        const double av = __builtin_isfinite(a[h]) ? a[h] : 0.0;
        image[i] = image[i] + fv * av;
        illum[i] = illum[i] + fv * fv;
    }
}

__global__ void finalize_kernel(int64_t n, const double* image, const double* illum, double* refl) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double il = illum[i];
        refl[i] = il > 1e-10 ? image[i] / il : 0.0;
    }
}

// g = -fwd * adj * (2 v dt dt), evaluated as the reference's numpy does:
// second_derivative = 2.0 * vel * dt * dt (left to right), then -fwd * adj * sd
__global__ void adjoint_gradient_kernel(int64_t n, const double* f, const double* a, const double* v, double dt,
                                        double* g) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double sd = ((2.0 * v[i]) * dt) * dt;
        g[i] = ((-f[i]) * a[i]) * sd;
    }
}

__global__ void accumulate_kernel(int64_t n, const float* x, double* y) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = y[i] + (double)x[i];
}

unsigned grid(int64_t n) {
    int64_t b = (n + 255) / 256;
    if (b > 8192) b = 8192;
    return (unsigned)(b < 1 ? 1 : b);
}

}  // namespace
}  // namespace stp

using stp::fail;
using stp::S;

extern "C" {

int stp_rtm_image_condition_f64(int64_t nx, int64_t ny, int64_t nz, int32_t hw, int64_t ny_stride, int64_t nz_stride,
                                const double* fwd, const double* adj, double* image, double* illum, void* stream) {
    if (nx < 0 || ny < 0 || nz < 0 || hw < 0 || !fwd || !adj || !image || !illum)
        return fail(STP_ERR_INVALID, "stp_rtm_image_condition_f64: bad argument");
    if (ny + 2 * hw > ny_stride || nz + 2 * hw > nz_stride)
        return fail(STP_ERR_INVALID, "stp_rtm_image_condition_f64: strides smaller than the with-halo extents");
    const int64_t n = nx * ny * nz;
    if (n == 0) return STP_OK;
    hipLaunchKernelGGL(stp::image_condition_kernel, dim3(stp::grid(n)), dim3(256), 0, S(stream), nx, ny, nz, hw,
                       ny_stride, nz_stride, fwd, adj, image, illum);
    STP_HIP(hipGetLastError());
    return STP_OK;
}

int stp_rtm_finalize_f64(int64_t n, const double* image, const double* illum, double* refl, void* stream) {
    if (n < 0 || (n > 0 && (!image || !illum || !refl))) return fail(STP_ERR_INVALID, "stp_rtm_finalize_f64: bad argument");
    if (n == 0) return STP_OK;
    hipLaunchKernelGGL(stp::finalize_kernel, dim3(stp::grid(n)), dim3(256), 0, S(stream), n, image, illum, refl);
    STP_HIP(hipGetLastError());
    return STP_OK;
}

int stp_adjoint_gradient_f64(int64_t n, const double* fwd, const double* adj, const double* vel, double dt,
                             double* grad, void* stream) {
    if (n < 0 || (n > 0 && (!fwd || !adj || !vel || !grad)))
        return fail(STP_ERR_INVALID, "stp_adjoint_gradient_f64: bad argument");
    if (n == 0) return STP_OK;
    hipLaunchKernelGGL(stp::adjoint_gradient_kernel, dim3(stp::grid(n)), dim3(256), 0, S(stream), n, fwd, adj, vel,
                       dt, grad);
    STP_HIP(hipGetLastError());
    return STP_OK;
}

int stp_accumulate_f32_f64(int64_t n, const float* x, double* y, void* stream) {
    if (n < 0 || (n > 0 && (!x || !y))) return fail(STP_ERR_INVALID, "stp_accumulate_f32_f64: bad argument");
    if (n == 0) return STP_OK;
    hipLaunchKernelGGL(stp::accumulate_kernel, dim3(stp::grid(n)), dim3(256), 0, S(stream), n, x, y);
    STP_HIP(hipGetLastError());
    return STP_OK;
}

}  // extern "C"
