// Travel-time tomography operator of the reference's Python inversion
// (src/python/seismic_tomography/inversion.py:329-433) on the GPU, FP64.
//
// A pick i joins a source and a station.  Both are mapped to cells by
// truncating (coord - origin) / h towards zero and clamping to [0, n-1]; the
// straight segment between the two points, of length dist, is sampled at
// n_s = max(int(dist / (5 dx)), 2) points frac = (s + 0.5) / n_s, each taken
// in the cell int(i_src + frac (i_sta - i_src)) (clamped), with weight
// ds = dist / n_s:
//   forward  t_i = sum_s ds / v(cell_s)      (v non-finite or <= 0 -> 1500;
//                                              t non-finite or < 0 -> dist / 6000)
//   gradient g[cell_s] += r_i (-ds / v^2) w_i (cells with an invalid v and
//                                              picks with a non-finite r_i skipped)
// One thread per pick.  The forward sum runs in the reference's order and
// without contracted multiply-adds, so it equals the reference's FP64 result
// bit for bit; the gradient is scattered with FP64 atomics (order of the
// additions per cell unspecified).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "capi_common.h"
#include "seismic_hip.h"

namespace {

#pragma clang fp contract(off)

struct Geometry {
    int64_t n[3];
    double o[3], h[3];
};

__device__ __forceinline__ int64_t cell_of(double c, double o, double h, int64_t n) {
    const double q = (c - o) / h;
    if (!(q < (double)(n - 1))) return q != q ? 0 : n - 1;   // NaN -> 0 (the reference raises)
    return q > 0.0 ? (int64_t)q : 0;
}

// sample s of pick i: flat cell index (x-major, z fastest)
struct Ray {
    int64_t a[3], b[3];
    double dist, ds;
    int64_t ns;
};

__device__ __forceinline__ Ray make_ray(const Geometry& g, const double* src, const double* sta) {
    Ray r;
    const double dx = src[0] - sta[0], dy = src[1] - sta[1], dz = src[2] - sta[2];
    r.dist = sqrt(dx * dx + dy * dy + dz * dz);
    for (int k = 0; k < 3; ++k) {
        r.a[k] = cell_of(src[k], g.o[k], g.h[k], g.n[k]);
        r.b[k] = cell_of(sta[k], g.o[k], g.h[k], g.n[k]);
    }
    // max(int(dist / (dx * 5)), 2); lengths beyond 2^62 samples (or NaN) are not
    // representable and take the minimum
    const double q = r.dist / (g.h[0] * 5.0);
    r.ns = (q >= 2.0 && q < 4.6e18) ? (int64_t)q : 2;
    r.ds = r.dist / (double)r.ns;
    return r;
}

__device__ __forceinline__ int64_t sample_cell(const Geometry& g, const Ray& r, int64_t s) {
    const double frac = ((double)s + 0.5) / (double)r.ns;
    int64_t c[3];
    for (int k = 0; k < 3; ++k) {
        const double p = (double)r.a[k] + frac * (double)(r.b[k] - r.a[k]);
        int64_t v = (int64_t)p;
        v = v < 0 ? 0 : (v > g.n[k] - 1 ? g.n[k] - 1 : v);
        c[k] = v;
    }
    return (c[0] * g.n[1] + c[1]) * g.n[2] + c[2];
}

__global__ void tt_forward_kernel(int64_t npicks, const double* __restrict__ src, const double* __restrict__ sta,
                                  Geometry g, const double* __restrict__ vp, double* __restrict__ t) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npicks) return;
    const Ray r = make_ray(g, src + 3 * i, sta + 3 * i);
    double tt = 0.0;
    for (int64_t s = 0; s < r.ns; ++s) {
        double v = vp[sample_cell(g, r, s)];
        if (!isfinite(v) || v <= 0.0) v = 1500.0;
        const double c = r.ds / v;
        if (!isfinite(c)) continue;
        tt += c;
    }
    if (!isfinite(tt) || tt < 0.0) tt = r.dist / 6000.0;
    t[i] = tt;
}

__global__ void tt_gradient_kernel(int64_t npicks, const double* __restrict__ src, const double* __restrict__ sta,
                                   Geometry g, const double* __restrict__ vp, const double* __restrict__ res,
                                   const double* __restrict__ w, double* grad) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npicks) return;
    const double ri = res[i];
    if (!isfinite(ri)) return;
    const double wi = w ? w[i] : 1.0;
    const Ray r = make_ray(g, src + 3 * i, sta + 3 * i);
    for (int64_t s = 0; s < r.ns; ++s) {
        const int64_t c = sample_cell(g, r, s);
        const double v = vp[c];
        if (!isfinite(v) || v <= 0.0) continue;
        const double f = -r.ds / (v * v);
        if (!isfinite(f)) continue;
        atomicAdd(grad + c, ri * f * wi);
    }
}

bool geometry(const int64_t* n, const double* origin, const double* h, Geometry* g) {
    if (!n || !origin || !h) return false;
    for (int k = 0; k < 3; ++k) {
        if (n[k] < 1 || !(h[k] > 0.0) || !std::isfinite(origin[k])) return false;
        g->n[k] = n[k];
        g->o[k] = origin[k];
        g->h[k] = h[k];
    }
    return g->n[0] * g->n[1] <= ((int64_t)1 << 62) / g->n[2];   // cell count fits int64
}

}  // namespace

extern "C" {

int stp_tt_forward(int64_t npicks, const double* src_xyz, const double* sta_xyz, const int64_t* n,
                   const double* origin, const double* h, const double* vp, double* times, void* stream) {
    Geometry g;
    if (npicks < 0 || !geometry(n, origin, h, &g) || (npicks > 0 && (!src_xyz || !sta_xyz || !vp || !times)))
        return stp::fail(STP_ERR_INVALID, "stp_tt_forward: bad argument");
    if (npicks == 0) return STP_OK;
    const int64_t nb = (npicks + 255) / 256;
    if (nb > 0x7fffffff) return stp::fail(STP_ERR_INVALID, "stp_tt_forward: too many picks");
    hipLaunchKernelGGL(tt_forward_kernel, dim3((unsigned)nb), dim3(256), 0, stp::S(stream), npicks, src_xyz, sta_xyz,
                       g, vp, times);
    STP_HIP(hipGetLastError());
    return STP_OK;
}

int stp_tt_gradient(int64_t npicks, const double* src_xyz, const double* sta_xyz, const int64_t* n,
                    const double* origin, const double* h, const double* vp, const double* weighted_residuals,
                    const double* weights, double* grad, void* stream) {
    Geometry g;
    if (npicks < 0 || !geometry(n, origin, h, &g) || !grad || !vp ||
        (npicks > 0 && (!src_xyz || !sta_xyz || !weighted_residuals)))
        return stp::fail(STP_ERR_INVALID, "stp_tt_gradient: bad argument");
    const int64_t V = g.n[0] * g.n[1] * g.n[2];
    STP_HIP(hipMemsetAsync(grad, 0, sizeof(double) * (size_t)V, stp::S(stream)));
    if (npicks == 0) return STP_OK;
    const int64_t nb = (npicks + 255) / 256;
    if (nb > 0x7fffffff) return stp::fail(STP_ERR_INVALID, "stp_tt_gradient: too many picks");
    hipLaunchKernelGGL(tt_gradient_kernel, dim3((unsigned)nb), dim3(256), 0, stp::S(stream), npicks, src_xyz,
                       sta_xyz, g, vp, weighted_residuals, weights, grad);
    STP_HIP(hipGetLastError());
    return STP_OK;
}

}  // extern "C"
