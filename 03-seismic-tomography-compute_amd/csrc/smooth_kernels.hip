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
    STP_HIP(r <= 8 ? stp::smooth_passes<8>(nx, ny, nz, r, wt, in, out, work, st)
                   : stp::smooth_passes<0>(nx, ny, nz, r, wt, in, out, work, st));
    return STP_OK;
}

}  // extern "C"
