// Gaussian smoothing of a gradient, SparseOps::smooth_gradient semantics
// (src/cpp/sparse/sparse_ops.cpp:115-158): weights exp(-|d|^2 / (2 sigma^2))
// over the cube |d_x|, |d_y|, |d_z| <= r = (int)(3 sigma), renormalised by the
// sum of the weights of the in-range cells.  The 3-D weight is a product of
// 1-D weights and the in-range set is a box, so the filter is exactly three
// 1-D passes (z, y, x), each divided by its own in-range weight sum; FP64.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

#include "capi_common.h"
#include "launch3.h"

namespace stp {
namespace {

constexpr int MAXR = 64;   // radius limit of the weight table (sigma < 21.67)

struct Weights {
    double w[MAXR + 1];   // 1-D weights for |d| = 0 .. r (passed by value)
};

// one pass along axis AX (0 x, 1 y, 2 z) of an nx x ny x nz volume; z from
// the thread, y and x from the block grid (launch3.h), so the coordinate
// along the axis is known without division.  RT > 0: r <= RT, and all 2 RT + 1
// tap loads are issued before the sum (a loop over a run-time radius waits
// for each load in turn); the in-range taps are added in the same order
// either way, so both forms give the same bits.
template <int AX, int ROWS, int RT>
__global__ void __launch_bounds__(256) smooth_pass_kernel(int64_t nx, int64_t ny, int64_t nz, int r, Weights wt,
                                                           const double* __restrict__ in, double* __restrict__ out) {
    __shared__ double ws[MAXR + 1];
    for (int k = threadIdx.x; k <= r; k += blockDim.x) ws[k] = wt.w[k];
    __syncthreads();
    const int64_t iz = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (iz >= nz) return;
    const int64_t stride = AX == 0 ? ny * nz : (AX == 1 ? nz : 1);
    const int64_t len = AX == 0 ? nx : (AX == 1 ? ny : nz);
    STP_FOR_PLANES_ROWS(ix, iy, nx, ny, ROWS) {
            const int64_t i = (ix * ny + iy) * nz + iz;
            const int64_t c = AX == 0 ? ix : (AX == 1 ? iy : iz);
            double s = 0.0, ws_sum = 0.0;
            if constexpr (RT > 0) {
                double val[2 * RT + 1];
                bool ok[2 * RT + 1];
#pragma unroll
                for (int t = 0; t < 2 * RT + 1; ++t) {
                    const int d = t - RT;
                    ok[t] = d >= -r && d <= r && c + d >= 0 && c + d < len;
                    const double x = in[i + (ok[t] ? d : 0) * stride];   // in bounds either way
                    val[t] = ok[t] ? x : 0.0;
                }
#pragma unroll
                for (int t = 0; t < 2 * RT + 1; ++t)
                    if (ok[t]) {
                        const double wd = ws[t < RT ? RT - t : t - RT];
                        s += wd * val[t];
                        ws_sum += wd;
                    }
            } else {
                for (int d = -r; d <= r; ++d) {
                    const int64_t j = c + d;
                    if (j < 0 || j >= len) continue;
                    const double wd = ws[d < 0 ? -d : d];
                    s += wd * in[i + d * stride];
                    ws_sum += wd;
                }
            }
            out[i] = s / ws_sum;
        }
}

// Fast forms for r <= RT, same taps added in the same order (same bits):
//
// x / y pass (AX = 0 / 1): each thread walks SEG outputs of one line along
// the axis with the 2 RT + 1 taps of the current output in registers plus PF
// values loaded ahead, so an output costs one load and the loads of PF later
// outputs are in flight.  Threads along z, blocks over the other coordinate
// (grid y) and over segments (grid z).
template <int AX, int RT, int SEG, int PF>
__global__ void __launch_bounds__(256) smooth_walk_kernel(int64_t nx, int64_t ny, int64_t nz, int r, Weights wt,
                                                           const double* __restrict__ in, double* __restrict__ out) {
    constexpr int NW = 2 * RT + 1 + PF;
    __shared__ double ws[MAXR + 1];
    for (int k = threadIdx.x; k <= r; k += blockDim.x) ws[k] = wt.w[k];
    __syncthreads();
    const int64_t iz = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (iz >= nz) return;
    const int64_t len = AX == 0 ? nx : ny, nother = AX == 0 ? ny : nx;
    const int64_t stride = AX == 0 ? ny * nz : nz;
    for (int64_t o = blockIdx.y; o < nother; o += gridDim.y) {
        const int64_t i0 = (AX == 0 ? o * nz : o * ny * nz) + iz;   // the line's element at axis coordinate 0
        auto ld = [&](int64_t j) {
            const bool ok = j >= 0 && j < len;
            const double x = in[i0 + (ok ? j : 0) * stride];   // in bounds either way
            return ok ? x : 0.0;
        };
        for (int64_t c0 = (int64_t)blockIdx.z * SEG; c0 < len; c0 += (int64_t)gridDim.z * SEG) {
            const int64_t c1 = c0 + SEG < len ? c0 + SEG : len;
            double win[NW];   // win[t] = value at c - RT + t
#pragma unroll
            for (int t = 0; t < NW; ++t) win[t] = ld(c0 - RT + t);
            for (int64_t c = c0; c < c1; ++c) {
                double s = 0.0, ws_sum = 0.0;
#pragma unroll
                for (int t = 0; t < 2 * RT + 1; ++t) {
                    const int d = t - RT;
                    if (d >= -r && d <= r && c + d >= 0 && c + d < len) {
                        const double wd = ws[t < RT ? RT - t : t - RT];
                        s += wd * win[t];
                        ws_sum += wd;
                    }
                }
                out[i0 + c * stride] = s / ws_sum;
                const double nxt = ld(c + RT + PF + 1);
#pragma unroll
                for (int t = 0; t < NW - 1; ++t) win[t] = win[t + 1];
                win[NW - 1] = nxt;
            }
        }
    }
}

// z pass: a block stages TZ outputs of one row plus RT halo values on each
// side in LDS and takes the taps from there; blocks walk `rows` rows.
template <int RT, int ROWS_>
__global__ void __launch_bounds__(256) smooth_z_lds_kernel(int64_t nx, int64_t ny, int64_t nz, int r, Weights wt,
                                                            const double* __restrict__ in, double* __restrict__ out) {
    constexpr int TZ = 256;
    __shared__ double ws[MAXR + 1];
    __shared__ double tile[TZ + 2 * RT];
    for (int k = threadIdx.x; k <= r; k += blockDim.x) ws[k] = wt.w[k];
    const int64_t z0 = (int64_t)blockIdx.x * TZ, iz = z0 + threadIdx.x;
    for (int64_t ix = blockIdx.z; ix < nx; ix += gridDim.z)
        for (int64_t y0 = (int64_t)blockIdx.y * ROWS_; y0 < ny; y0 += (int64_t)gridDim.y * ROWS_)
            for (int64_t iy = y0, y1 = (y0 + ROWS_ < ny ? y0 + ROWS_ : ny); iy < y1; ++iy) {
                const int64_t row = (ix * ny + iy) * nz;
                __syncthreads();   // the previous row's taps are read
                for (int k = threadIdx.x; k < TZ + 2 * RT; k += blockDim.x) {
                    const int64_t z = z0 - RT + k;
                    tile[k] = (z >= 0 && z < nz) ? in[row + z] : 0.0;
                }
                __syncthreads();
                if (iz < nz) {
                    double s = 0.0, ws_sum = 0.0;
#pragma unroll
                    for (int t = 0; t < 2 * RT + 1; ++t) {
                        const int d = t - RT;
                        if (d >= -r && d <= r && iz + d >= 0 && iz + d < nz) {
                            const double wd = ws[t < RT ? RT - t : t - RT];
                            s += wd * tile[threadIdx.x + t];
                            ws_sum += wd;
                        }
                    }
                    out[row + iz] = s / ws_sum;
                }
            }
}

inline dim3 walk_grid(int64_t nz, int64_t nother, int64_t len, int seg) {
    const int64_t bz = (nz + 255) / 256, nseg = (len + seg - 1) / seg;
    return dim3((unsigned)(bz < 1 ? 1 : bz), (unsigned)(nother < 1 ? 1 : (nother > 65535 ? 65535 : nother)),
                (unsigned)(nseg < 1 ? 1 : (nseg > 65535 ? 65535 : nseg)));
}

hipError_t smooth_passes_fast(int64_t nx, int64_t ny, int64_t nz, int r, const Weights& wt, const double* in,
                              double* out, double* work, hipStream_t st) {
    constexpr int RT = 8, SEG = 64, PF = 4;
    hipLaunchKernelGGL((smooth_z_lds_kernel<RT, ROWS>), grid3(nx, ny, nz), dim3(256), 0, st, nx, ny, nz, r, wt, in,
                       out);
    hipLaunchKernelGGL((smooth_walk_kernel<1, RT, SEG, PF>), walk_grid(nz, nx, ny, SEG), dim3(256), 0, st, nx, ny, nz,
                       r, wt, out, work);
    hipLaunchKernelGGL((smooth_walk_kernel<0, RT, SEG, PF>), walk_grid(nz, ny, nx, SEG), dim3(256), 0, st, nx, ny, nz,
                       r, wt, work, out);
    return hipGetLastError();
}

template <int RT>
hipError_t smooth_passes(int64_t nx, int64_t ny, int64_t nz, int r, const Weights& wt, const double* in,
                         double* out, double* work, hipStream_t st) {
    // the y pass reads 2r neighbouring rows, so its blocks take longer chunks
    constexpr int RY = 64;
    const dim3 g3 = grid3(nx, ny, nz), g3y = grid3(nx, ny, nz, RY);
    hipLaunchKernelGGL((smooth_pass_kernel<2, ROWS, RT>), g3, dim3(256), 0, st, nx, ny, nz, r, wt, in, out);
    hipLaunchKernelGGL((smooth_pass_kernel<1, RY, RT>), g3y, dim3(256), 0, st, nx, ny, nz, r, wt, out, work);
    hipLaunchKernelGGL((smooth_pass_kernel<0, ROWS, RT>), g3, dim3(256), 0, st, nx, ny, nz, r, wt, work, out);
    return hipGetLastError();
}

// SparseOps::gaussian_filter_3d (sparse_ops.cpp:212-234): the Gaussian
// exp(-|c|^2 / (2 sigma^2)) of every cell, c = (ix, iy, iz) - (nx, ny, nz) / 2.0,
// normalised by the sum over the grid.  Pass 1 writes the values and one
// partial sum per block; pass 2 has every block add the partials in the same
// order and scale its cells.
constexpr int GAUSS_BLOCKS = 1024;

__global__ void __launch_bounds__(256) gauss3_values_kernel(int64_t nx, int64_t ny, int64_t nz, double s2,
                                                            double* __restrict__ out, double* __restrict__ part) {
    const int64_t n = nx * ny * nz;
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t ix = i / (ny * nz), iy = (i / nz) % ny, iz = i % nz;
        const double cx = (double)ix - (double)nx / 2.0, cy = (double)iy - (double)ny / 2.0,
                     cz = (double)iz - (double)nz / 2.0;
        const double v = exp(-(cx * cx + cy * cy + cz * cz) / s2);
        out[i] = v;
        acc += v;
    }
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) gauss3_scale_kernel(int64_t n, double* __restrict__ out,
                                                           const double* __restrict__ part, int np) {
    __shared__ double red[256];
    double acc = 0.0;
    for (int k = threadIdx.x; k < np; k += 256) acc += part[k];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    const double inv = 1.0 / red[0];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] *= inv;
}

}  // namespace
}  // namespace stp

using stp::fail;
using stp::S;

extern "C" {

int stp_smooth_gradient_f64(int64_t nx, int64_t ny, int64_t nz, double sigma, const double* in, double* out,
                            double* work, void* stream) {
    if (nx < 0 || ny < 0 || nz < 0 || !(sigma > 0.0) || !std::isfinite(sigma))
        return fail(STP_ERR_INVALID, "stp_smooth_gradient_f64: bad argument");
    const int64_t n = nx * ny * nz;
    if (n == 0) return STP_OK;
    if (!in || !out || !work) return fail(STP_ERR_INVALID, "stp_smooth_gradient_f64: null pointer");
    const double r3 = 3.0 * sigma;
    if (r3 >= stp::MAXR + 1) return fail(STP_ERR_INVALID, "stp_smooth_gradient_f64: sigma too large (radius > 64)");
    const int r = (int)r3;
    stp::Weights wt{};
    for (int d = 0; d <= r; ++d) wt.w[d] = std::exp(-(double)(d * d) / (2.0 * sigma * sigma));
    hipStream_t st = S(stream);
    STP_HIP(r <= 8 ? stp::smooth_passes_fast(nx, ny, nz, r, wt, in, out, work, st)
                   : stp::smooth_passes<0>(nx, ny, nz, r, wt, in, out, work, st));
    return STP_OK;
}

int stp_gaussian_filter_3d_f64(int64_t nx, int64_t ny, int64_t nz, double sigma, double* out, double* work,
                               void* stream) {
    if (nx < 0 || ny < 0 || nz < 0 || !(sigma > 0.0) || !std::isfinite(sigma))
        return fail(STP_ERR_INVALID, "stp_gaussian_filter_3d_f64: bad argument");
    const int64_t n = nx * ny * nz;
    if (n == 0) return STP_OK;
    if (!out || !work) return fail(STP_ERR_INVALID, "stp_gaussian_filter_3d_f64: null pointer");
    int64_t nb = (n + 255) / 256;
    if (nb > stp::GAUSS_BLOCKS) nb = stp::GAUSS_BLOCKS;
    hipStream_t st = S(stream);
    hipLaunchKernelGGL(stp::gauss3_values_kernel, dim3((unsigned)nb), dim3(256), 0, st, nx, ny, nz,
                       2.0 * sigma * sigma, out, work);
    STP_HIP(hipGetLastError());
    hipLaunchKernelGGL(stp::gauss3_scale_kernel, dim3((unsigned)nb), dim3(256), 0, st, n, out, work, (int)nb);
    STP_HIP(hipGetLastError());
    return STP_OK;
}

}  // extern "C"
