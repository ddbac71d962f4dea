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

namespace stp {
namespace {

constexpr int MAXR = 64;   // radius limit of the weight table (sigma < 21.67)

struct Weights {
    double w[MAXR + 1];   // 1-D weights for |d| = 0 .. r (passed by value)
};

// one pass along the axis of stride `stride` and length `len`
__global__ void smooth_pass_kernel(int64_t n, int64_t len, int64_t stride, int r, Weights wt,
                                   const double* __restrict__ in, double* __restrict__ out) {
    __shared__ double ws[MAXR + 1];
    for (int k = threadIdx.x; k <= r; k += blockDim.x) ws[k] = wt.w[k];
    __syncthreads();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = (i / stride) % len;
        double s = 0.0, ws_sum = 0.0;
        for (int d = -r; d <= r; ++d) {
            const int64_t j = c + d;
            if (j < 0 || j >= len) continue;
            const double wd = ws[d < 0 ? -d : d];
            s += wd * in[i + d * stride];
            ws_sum += wd;
        }
        out[i] = s / ws_sum;
    }
}

unsigned grid(int64_t n) {
    int64_t b = (n + 255) / 256;
    if (b > 16384) b = 16384;
    return (unsigned)(b < 1 ? 1 : b);
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
    hipLaunchKernelGGL(stp::smooth_pass_kernel, dim3(stp::grid(n)), dim3(256), 0, st, n, nz, (int64_t)1, r, wt, in,
                       out);
    hipLaunchKernelGGL(stp::smooth_pass_kernel, dim3(stp::grid(n)), dim3(256), 0, st, n, ny, nz, r, wt, out, work);
    hipLaunchKernelGGL(stp::smooth_pass_kernel, dim3(stp::grid(n)), dim3(256), 0, st, n, nx, ny * nz, r, wt, work,
                       out);
    STP_HIP(hipGetLastError());
    return STP_OK;
}

}  // extern "C"
