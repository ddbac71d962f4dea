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

// ------------------------------------------------------------------ slice capture
// RTMEngine::capture_single_slice (src/cpp/rtm/rtm_engine.cpp:83-165): one
// plane of the forward and adjoint with-halo fields (strides ny_s * nz_s and
// nz_s, the caller's) and of the image condition, non-finite values -> 0.
// The image condition is indexed as the reference indexes it: by the two
// in-plane coordinates only (axis X: ic[iy*nz + iz], Y: ic[ix*nz + iz], Z:
// ic[ix*ny + iy]).
namespace {
__global__ void extract_slice_kernel(int axis, int64_t idx, int64_t nx, int64_t ny, int64_t nz, int64_t hw,
                                     int64_t ny_s, int64_t nz_s, const double* fwd, const double* adj,
                                     const double* ic, double* of, double* oa, double* oi) {
    const int64_t d0 = axis == 0 ? ny : nx, d1 = axis == 2 ? ny : nz;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < d0 * d1;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t a = t / d1, b = t % d1;
        int64_t src, ici;
        if (axis == 0) {
            src = (idx + hw) * (ny_s * nz_s) + (a + hw) * nz_s + (b + hw);
            ici = a * nz + b;
        } else if (axis == 1) {
            src = (a + hw) * (ny_s * nz_s) + (idx + hw) * nz_s + (b + hw);
            ici = a * nz + b;
        } else {
            src = (a + hw) * (ny_s * nz_s) + (b + hw) * nz_s + (idx + hw);
            ici = a * ny + b;
        }
        const double f = fwd[src], g = adj[src], c = ic[ici];
        of[t] = isfinite(f) ? f : 0.0;
        oa[t] = isfinite(g) ? g : 0.0;
        oi[t] = isfinite(c) ? c : 0.0;
    }
}
}  // namespace

extern "C" int stp_rtm_extract_slice_f64(int32_t axis, int64_t index, int64_t nx, int64_t ny, int64_t nz,
                                         int32_t halo_width, int64_t ny_stride, int64_t nz_stride,
                                         const double* fwd, const double* adj, const double* ic, double* out_fwd,
                                         double* out_adj, double* out_ic, void* stream) {
    if (axis < 0 || axis > 2 || nx < 1 || ny < 1 || nz < 1 || halo_width < 0 || !fwd || !adj || !ic || !out_fwd ||
        !out_adj || !out_ic)
        return stp::fail(STP_ERR_INVALID, "stp_rtm_extract_slice_f64: bad argument");
    const int64_t ext = axis == 0 ? nx : (axis == 1 ? ny : nz);
    if (index < 0 || index >= ext) return stp::fail(STP_ERR_INVALID, "stp_rtm_extract_slice_f64: index out of range");
    const int64_t n = (axis == 0 ? ny : nx) * (axis == 2 ? ny : nz);
    int64_t nb = (n + 255) / 256;
    nb = nb < 1 ? 1 : (nb > 65535 ? 65535 : nb);
    hipLaunchKernelGGL(extract_slice_kernel, dim3((unsigned)nb), dim3(256), 0, stp::S(stream), (int)axis, index, nx, ny,
                       nz, (int64_t)halo_width, ny_stride, nz_stride, fwd, adj, ic, out_fwd, out_adj, out_ic);
    STP_HIP(hipGetLastError());
    return STP_OK;
}
