// Sparse kernels for the tomography solve on gfx950.
//
//   * CSR SpMV: W lanes of a wavefront cooperate on one row (W = 16/32/64
//     chosen from the mean row length), lane-strided coalesced loads of the
//     FP32 values and int32 columns, gather of x through L2, reduction with
//     cross-lane shuffles.  Optional fused epilogues: y = a*Ax - lam*w and
//     per-block FP64 partial sums of y*y / y*z for the solver's dot products.
//   * Siddon straight-ray matrix: one thread per ray, exact path lengths in
//     FP64 (no FP contraction, so they equal the CPU restatement bit for bit),
//     rounded to FP32; rows sorted by column.
//   * Transpose: radix sort of (col, row) keys — deterministic order.
//   * Solver vector kernels with device-resident scalars: no host round trip
//     per iteration; a `done` flag lets kernels of iterations queued after
//     convergence return immediately.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>
#include <stdlib.h>

#include "sparse_kernels.h"

namespace stp {

// ------------------------------------------------------------------ SpMV
// Rows are processed by groups of W lanes (RPB = 256 / W rows per block per
// pass); blocks stride over the rows, so the grid is bounded (<= MAX_BLOCKS)
// and each block leaves one FP64 partial per reduced quantity.
constexpr int MAX_BLOCKS = SPMV_MAX_BLOCKS;

// Each lane takes entries b+lane, b+lane+W, ... and accumulates them in that
// order.  The row is consumed in chunks of U*W entries (U*W = 128): all U
// value/column loads of a chunk are issued, then all U gathers of x, then
// the U products are added, so a lane has U independent gathers in flight
// instead of a chain of dependent load pairs.  Slots past the row end load
// the row's first entry and contribute an exact zero, which leaves the sum
// (and its order) unchanged.
template <int W, typename AT, typename XT, typename VT = float, bool NOVAL = false, typename IT = int>
__device__ __forceinline__ AT row_dot_range(int b, int e, const IT* __restrict__ ci, const VT* __restrict__ val,
                                            const XT* __restrict__ x, int lane) {
    constexpr int U = (128 / W) < 1 ? 1 : (128 / W);
    AT acc = 0;
    for (int k0 = b + lane; k0 < e; k0 += U * W) {
        VT v[U];
        int c[U];
        bool ok[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = k0 + u * W;
            ok[u] = k < e;
            const int kk = ok[u] ? k : b;
            if constexpr (NOVAL) v[u] = (VT)1;
            else v[u] = val[kk];
            c[u] = ci[kk];
        }
        AT xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) xv[u] = (AT)x[c[u]];
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (ok[u]) acc += (AT)v[u] * xv[u];
    }
#pragma unroll
    for (int m = W / 2; m > 0; m >>= 1) acc += __shfl_xor(acc, m);
    return acc;
}

// The same for row `row` of the CSR arrays (rp, ci, val).
template <int W, typename AT, typename XT, typename VT = float, bool NOVAL = false, typename IT = int>
__device__ __forceinline__ AT row_dot(const int* __restrict__ rp, const IT* __restrict__ ci,
                                      const VT* __restrict__ val, const XT* __restrict__ x, int row, int lane) {
    return row_dot_range<W, AT, XT, VT, NOVAL, IT>(rp[row], rp[row + 1], ci, val, x, lane);
}

// Block reductions for 256-thread blocks (4 wavefronts): a butterfly of
// cross-lane adds inside each wavefront (every lane ends with the same sum),
// then the four wavefront sums added in wavefront order.  The order is fixed,
// so the sums are reproducible.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m);
    return v;
}

__device__ __forceinline__ void block_partial(double v, double* part) {
    __shared__ double red[4];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
    __syncthreads();   // red is reused by the next call
}

template <int W, typename XT, typename AT, typename YT, bool FUSE_W, bool PART, typename VT = float>
__global__ void __launch_bounds__(256) spmv_kernel(int rows, const int* __restrict__ rp, const int* __restrict__ ci,
                                                    const VT* __restrict__ val, const XT* __restrict__ x,
                                                    YT* __restrict__ y, double alpha, const double* __restrict__ wv,
                                                    double lam, double* __restrict__ part, const int* done,
                                                    int rev) {
    if (done && *done) return;
    constexpr int RPB = 256 / W;
    const int lane = threadIdx.x % W, sub = threadIdx.x / W;
    double ps = 0.0;
    // Offsets of the group's rows are loaded one row ahead (a group past the
    // last row loads the last row's and does not use them).
    auto offsets = [&](int row0, int& b, int& e) {
        const int k = row0 < rows - sub ? row0 + sub : rows - 1;
        const int row = rev ? rows - 1 - k : k;
        b = rp[row];
        e = rp[row + 1];
    };
    int b = 0, e = 0;
    if (blockIdx.x * RPB < rows) offsets(blockIdx.x * RPB, b, e);
    for (int row0 = blockIdx.x * RPB; row0 < rows; row0 += gridDim.x * RPB) {
        int nb, ne;
        offsets(row0 < rows - (int)gridDim.x * RPB ? row0 + (int)gridDim.x * RPB : rows - 1, nb, ne);
        if (row0 + sub < rows) {
            const int row = rev ? rows - 1 - (row0 + sub) : row0 + sub;
            const AT acc = row_dot_range<W, AT, XT, VT>(b, e, ci, val, x, lane);
            if (lane == 0) {
                double v = alpha * (double)acc;
                if (FUSE_W) v -= lam * wv[row];
                const YT yv = (YT)v;
                y[row] = yv;
                if (PART) ps += (double)yv * (double)yv;
            }
        }
        b = nb;
        e = ne;
    }
    if (PART) block_partial(ps, part);
}

template <int W, typename XT, typename AT, typename YT, bool FUSE_W, bool PART, typename VT = float>
static hipError_t spmv_w(int rows, const int* rp, const int* ci, const VT* val, const XT* x, YT* y, double alpha,
                         const double* wv, double lam, double* part, const int* done, hipStream_t st, int rev = 0) {
    const int nb = spmv_blocks(rows, W);
    if (rows == 0) return hipSuccess;
    hipLaunchKernelGGL((spmv_kernel<W, XT, AT, YT, FUSE_W, PART, VT>), dim3(nb), dim3(256), 0, st, rows, rp, ci, val,
                       x, y, alpha, wv, lam, part, done, rev);
    return hipGetLastError();
}

// Lanes per row.  A (rays x cells, FP32 gathers of a cell vector) and A^T
// (cells x rays, FP64 gathers of a ray vector) are chosen separately: at
// 200k rays x 64^3 cells (107 and 81 entries per row) A runs fastest with
// 32 lanes and A^T with 64 (a whole wavefront gathers the rays of one cell,
// which are close in ray order).  STP_SPMV_W / STP_SPMVT_W override.
int spmv_lanes(double mean_nnz, int transpose) {
    if (const char* ev = getenv(transpose ? "STP_SPMVT_W" : "STP_SPMV_W")) {
        const int w = atoi(ev);
        if (w == 8 || w == 16 || w == 32 || w == 64) return w;
    }
    if (transpose && mean_nnz >= 48) return 64;
    return mean_nnz >= 96 ? 32 : (mean_nnz >= 24 ? 16 : 8);
}

// Lanes per row in the solver layout.  Consecutive rows there are neighbouring
// cells (A^T) or neighbouring rays (A), so a wavefront that takes several
// rows at once shares gathered lines between them: at 200k rays x 64^3 cells
// A^T runs fastest with 8 lanes and A with 16 (a CGLS iteration 174 us,
// against 237 us with the widths of the given order).  STP_SPMVL_W /
// STP_SPMVLT_W override.
int spmv_lanes_layout(double mean_nnz, int transpose) {
    if (const char* ev = getenv(transpose ? "STP_SPMVLT_W" : "STP_SPMVL_W")) {
        const int w = atoi(ev);
        if (w == 8 || w == 16 || w == 32 || w == 64) return w;
    }
    if (transpose) return 8;
    return mean_nnz >= 24 ? 16 : 8;
}

int spmv_blocks(int rows, int W) {
    const int nb = (rows + 256 / W - 1) / (256 / W);
    return nb < 1 ? 1 : (nb > MAX_BLOCKS ? MAX_BLOCKS : nb);
}

#define STP_SPMV_DISPATCH(XT, AT, YT, FW, PT)                                                                   \
    switch (W) {                                                                                                \
        case 64: return spmv_w<64, XT, AT, YT, FW, PT>(rows, rp, ci, val, x, y, alpha, wv, lam, part, done, st, rev); \
        case 32: return spmv_w<32, XT, AT, YT, FW, PT>(rows, rp, ci, val, x, y, alpha, wv, lam, part, done, st, rev); \
        case 16: return spmv_w<16, XT, AT, YT, FW, PT>(rows, rp, ci, val, x, y, alpha, wv, lam, part, done, st, rev); \
        default: return spmv_w<8, XT, AT, YT, FW, PT>(rows, rp, ci, val, x, y, alpha, wv, lam, part, done, st, rev);  \
    }

hipError_t spmv_f32(int W, int rows, const int* rp, const int* ci, const float* val, const float* x, float* y,
                    double* part, const int* done, hipStream_t st, int rev) {
    const double alpha = 1.0, lam = 0.0;
    const double* wv = nullptr;
    if (part) { STP_SPMV_DISPATCH(float, float, float, false, true) }
    STP_SPMV_DISPATCH(float, float, float, false, false)
}

hipError_t spmv_f64(int W, int rows, const int* rp, const int* ci, const float* val, const double* x, double* y,
                    double alpha, const double* wv, double lam, double* part, const int* done, hipStream_t st) {
    const int rev = 0;
    if (wv) {
        if (part) { STP_SPMV_DISPATCH(double, double, double, true, true) }
        STP_SPMV_DISPATCH(double, double, double, true, false)
    }
    if (part) { STP_SPMV_DISPATCH(double, double, double, false, true) }
    STP_SPMV_DISPATCH(double, double, double, false, false)
}

// FP64 values, FP64 x and accumulation: the square CG of the reference's
// all-FP64 arithmetic (12 bytes per nonzero instead of 8).
hipError_t spmv_f64v(int W, int rows, const int* rp, const int* ci, const double* val, const double* x, double* y,
                     const int* done, hipStream_t st) {
    const double alpha = 1.0, lam = 0.0;
    const double* wv = nullptr;
    double* part = nullptr;
    switch (W) {
        case 64: return spmv_w<64, double, double, double, false, false, double>(rows, rp, ci, val, x, y, alpha, wv,
                                                                                  lam, part, done, st);
        case 32: return spmv_w<32, double, double, double, false, false, double>(rows, rp, ci, val, x, y, alpha, wv,
                                                                                  lam, part, done, st);
        case 16: return spmv_w<16, double, double, double, false, false, double>(rows, rp, ci, val, x, y, alpha, wv,
                                                                                  lam, part, done, st);
        default: return spmv_w<8, double, double, double, false, false, double>(rows, rp, ci, val, x, y, alpha, wv,
                                                                                 lam, part, done, st);
    }
}

// Gather probe: the SpMV kernel without its value stream (row sums of the
// gathered x[ci[k]], same lanes per row, chunking and row order), i.e. the
// index stream plus the gathers of one product.  Measurement only.
template <int W, typename XT>
__global__ void __launch_bounds__(256) gather_probe_kernel(int rows, const int* __restrict__ rp,
                                                           const int* __restrict__ ci, const XT* __restrict__ x,
                                                           XT* __restrict__ y) {
    constexpr int RPB = 256 / W;
    const int lane = threadIdx.x % W, sub = threadIdx.x / W;
    for (int row0 = blockIdx.x * RPB; row0 < rows; row0 += gridDim.x * RPB) {
        const int row = row0 + sub;
        if (row < rows) {
            const XT acc = row_dot<W, XT, XT, float, true>(rp, ci, nullptr, x, row, lane);
            if (lane == 0) y[row] = acc;
        }
    }
}

template <typename XT>
hipError_t gather_probe(int W, int rows, const int* rp, const int* ci, const XT* x, XT* y, hipStream_t st) {
    if (rows == 0) return hipSuccess;
    const int nb = spmv_blocks(rows, W);
    switch (W) {
        case 64: hipLaunchKernelGGL((gather_probe_kernel<64, XT>), dim3(nb), dim3(256), 0, st, rows, rp, ci, x, y); break;
        case 32: hipLaunchKernelGGL((gather_probe_kernel<32, XT>), dim3(nb), dim3(256), 0, st, rows, rp, ci, x, y); break;
        case 16: hipLaunchKernelGGL((gather_probe_kernel<16, XT>), dim3(nb), dim3(256), 0, st, rows, rp, ci, x, y); break;
        default: hipLaunchKernelGGL((gather_probe_kernel<8, XT>), dim3(nb), dim3(256), 0, st, rows, rp, ci, x, y); break;
    }
    return hipGetLastError();
}
template hipError_t gather_probe<float>(int, int, const int*, const int*, const float*, float*, hipStream_t);
template hipError_t gather_probe<double>(int, int, const int*, const int*, const double*, double*, hipStream_t);

// s = A^T r - lam x with FP64 accumulation, FP32 output s, z = dinv*s,
// partials of s*z and s*s.  r is gathered as RT: FP64 for the true residual,
// FP32 inside the CGLS iteration.  With gin the gathered vector is the change
// y = r_ref - r of the residual since the pass started and gin = A^T r_ref,
// so s = gin - A^T y - lam x; gout (when given) receives A^T r in FP64.
// rev != 0 visits the rows last to first.
template <int W, typename RT>
__global__ void __launch_bounds__(256) normal_residual_kernel(int rows, const int* __restrict__ rp,
                                                               const int* __restrict__ ci, const float* __restrict__ val,
                                                               const RT* __restrict__ r,
                                                               const double* __restrict__ x, double lam,
                                                               const float* __restrict__ dinv, float* __restrict__ s,
                                                               float* __restrict__ z, double* __restrict__ part_sz,
                                                               double* __restrict__ part_ss, const int* done, int rev,
                                                               const double* __restrict__ gin,
                                                               double* __restrict__ gout) {
    if (done && *done) return;
    constexpr int RPB = 256 / W;
    const int lane = threadIdx.x % W, sub = threadIdx.x / W;
    double psz = 0.0, pss = 0.0;
    // row offsets one row ahead, as in spmv_kernel
    auto offsets = [&](int row0, int& b, int& e) {
        const int k = row0 < rows - sub ? row0 + sub : rows - 1;
        const int row = rev ? rows - 1 - k : k;
        b = rp[row];
        e = rp[row + 1];
    };
    int b = 0, e = 0;
    if (blockIdx.x * RPB < rows) offsets(blockIdx.x * RPB, b, e);
    for (int row0 = blockIdx.x * RPB; row0 < rows; row0 += gridDim.x * RPB) {
        int nb, ne;
        offsets(row0 < rows - (int)gridDim.x * RPB ? row0 + (int)gridDim.x * RPB : rows - 1, nb, ne);
        if (row0 + sub < rows) {
            const int row = rev ? rows - 1 - (row0 + sub) : row0 + sub;
            const double acc = row_dot_range<W, double, RT>(b, e, ci, val, r, lane);
            if (lane == 0) {
                if (gout) gout[row] = acc;
                const float sv = (float)((gin ? gin[row] - acc : acc) - lam * x[row]);
                const float zv = dinv ? dinv[row] * sv : sv;
                s[row] = sv;
                z[row] = zv;
                psz += (double)sv * (double)zv;
                pss += (double)sv * (double)sv;
            }
        }
        b = nb;
        e = ne;
    }
    block_partial(psz, part_sz);
    block_partial(pss, part_ss);
}

template <typename RT>
static hipError_t normal_residual_t(int W, int rows, const int* rp, const int* ci, const float* val, const RT* r,
                                    const double* x, double lam, const float* dinv, float* s, float* z,
                                    double* part_sz, double* part_ss, const int* done, int rev, const double* gin,
                                    double* gout, hipStream_t st) {
    const int nb = spmv_blocks(rows, W);
    if (rows == 0) return hipSuccess;
    switch (W) {
        case 64: hipLaunchKernelGGL((normal_residual_kernel<64, RT>), dim3(nb), dim3(256), 0, st, rows, rp, ci, val, r,
                                    x, lam, dinv, s, z, part_sz, part_ss, done, rev, gin, gout); break;
        case 32: hipLaunchKernelGGL((normal_residual_kernel<32, RT>), dim3(nb), dim3(256), 0, st, rows, rp, ci, val, r,
                                    x, lam, dinv, s, z, part_sz, part_ss, done, rev, gin, gout); break;
        case 16: hipLaunchKernelGGL((normal_residual_kernel<16, RT>), dim3(nb), dim3(256), 0, st, rows, rp, ci, val, r,
                                    x, lam, dinv, s, z, part_sz, part_ss, done, rev, gin, gout); break;
        default: hipLaunchKernelGGL((normal_residual_kernel<8, RT>), dim3(nb), dim3(256), 0, st, rows, rp, ci, val, r,
                                    x, lam, dinv, s, z, part_sz, part_ss, done, rev, gin, gout); break;
    }
    return hipGetLastError();
}

hipError_t normal_residual(int W, int rows, const int* rp, const int* ci, const float* val, const double* r,
                           const double* x, double lam, const float* dinv, float* s, float* z, double* part_sz,
                           double* part_ss, const int* done, hipStream_t st, int rev, double* gout) {
    return normal_residual_t<double>(W, rows, rp, ci, val, r, x, lam, dinv, s, z, part_sz, part_ss, done, rev,
                                     nullptr, gout, st);
}

hipError_t normal_residual_r32(int W, int rows, const int* rp, const int* ci, const float* val, const float* y,
                               const double* g, const double* x, double lam, const float* dinv, float* s, float* z,
                               double* part_sz, double* part_ss, const int* done, int rev, hipStream_t st) {
    return normal_residual_t<float>(W, rows, rp, ci, val, y, x, lam, dinv, s, z, part_sz, part_ss, done, rev, g,
                                    nullptr, st);
}

// Staged A^T y: one workgroup per tile of rt rows (cells) at a time.  The
// tile's distinct rays are read from y once, into LDS (uniq is sorted, so
// neighbouring lanes read neighbouring rays), and the rows' dot products then
// gather from LDS through 16-bit positions; the row order, lanes and chunking
// are those of normal_residual_kernel, so each row's sum is the same.  As
// there, with g the gathered vector is y and s = g - A^T y - lam x.  HG / HD:
// g / the Jacobi scaling dinv are given (compile-time, so that no load sits
// under a run-time condition that would make the compiler wait for it early).
template <int W, bool HG, bool HD>
__global__ void __launch_bounds__(256) normal_residual_staged_kernel(
    int rows, int rt, int ntiles, const int* __restrict__ rp, const uint16_t* __restrict__ l16,
    const float* __restrict__ val, const int* __restrict__ tu, const int* __restrict__ uniq,
    const float* __restrict__ r, const double* __restrict__ g, const double* __restrict__ x, double lam,
    const float* __restrict__ dinv, float* __restrict__ s, float* __restrict__ z, double* __restrict__ part_sz,
    double* __restrict__ part_ss, const int* done, int rev) {
    extern __shared__ float xs[];
    if (done && *done) return;
    constexpr int RPB = 256 / W;
    const int lane = threadIdx.x % W, sub = threadIdx.x / W;
    double psz = 0.0, pss = 0.0;
    for (int t0 = blockIdx.x; t0 < ntiles; t0 += gridDim.x) {
        const int t = rev ? ntiles - 1 - t0 : t0;
        const int ub = tu[t], nu = tu[t + 1] - ub;
        const int r1 = min(t * rt + rt, rows);
        // Row offsets of the group's first row (rows t rt .. r1 - 1 are not
        // empty), loaded with the staging loads.
        int b, e;
        {
            const int rf = min(t * rt + sub, r1 - 1);
            b = rp[rf];
            e = rp[rf + 1];
        }
        __syncthreads();   // the previous tile's reads of xs are done
        // Stage the tile's rays: SU list entries per thread are loaded
        // together, then their gathers, so a thread waits two memory latencies
        // per 256 SU rays instead of two per ray.  Slots past the list's end
        // take its last entry and store the same value to its slot (no
        // branch, so the compiler keeps every load in the batch).
        constexpr int SU = 8;
        for (int k0 = threadIdx.x; k0 < nu; k0 += SU * 256) {
            int id[SU];
            float rv[SU];
#pragma unroll
            for (int u = 0; u < SU; ++u) id[u] = uniq[ub + min(k0 + u * 256, nu - 1)];
#pragma unroll
            for (int u = 0; u < SU; ++u) rv[u] = r[id[u]];
#pragma unroll
            for (int u = 0; u < SU; ++u) xs[min(k0 + u * 256, nu - 1)] = rv[u];
        }
        __syncthreads();
        for (int row = t * rt + sub; row < r1; row += RPB) {
            // the next row's offsets and this row's epilogue operands go out
            // before the row's own loads
            const int rn = min(row + RPB, r1 - 1);
            const int nb = rp[rn], ne = rp[rn + 1];
            const double gr = HG ? g[row] : 0.0, xr = x[row];
            const float dr = HD ? dinv[row] : 1.0f;
            const double acc = row_dot_range<W, double, float, float, false, uint16_t>(b, e, l16, val, xs, lane);
            b = nb;
            e = ne;
            if (lane == 0) {
                const float sv = (float)((HG ? gr - acc : acc) - lam * xr);
                const float zv = HD ? dr * sv : sv;
                s[row] = sv;
                z[row] = zv;
                psz += (double)sv * (double)zv;
                pss += (double)sv * (double)sv;
            }
        }
    }
    block_partial(psz, part_sz);
    block_partial(pss, part_ss);
}

int staged_blocks(int ntiles) { return ntiles < 1 ? 1 : (ntiles > MAX_BLOCKS ? MAX_BLOCKS : ntiles); }

hipError_t normal_residual_staged(int W, int rows, int rt, int ntiles, int maxu, const int* rp, const uint16_t* l16,
                                  const float* val, const int* tu, const int* uniq, const float* y, const double* g,
                                  const double* x, double lam, const float* dinv, float* s, float* z,
                                  double* part_sz, double* part_ss, const int* done, int rev, hipStream_t st) {
    if (rows == 0) return hipSuccess;
    if (maxu < 0 || maxu > STAGE_CAP || rt < 1) return hipErrorInvalidValue;
    const int nb = staged_blocks(ntiles);
    const size_t lds = sizeof(float) * (size_t)(maxu > 0 ? maxu : 1);
#define STP_STAGED(WW)                                                                                          \
    do {                                                                                                        \
        if (g && dinv) STP_STAGED_LAUNCH(WW, true, true);                                                       \
        else if (g) STP_STAGED_LAUNCH(WW, true, false);                                                         \
        else if (dinv) STP_STAGED_LAUNCH(WW, false, true);                                                      \
        else STP_STAGED_LAUNCH(WW, false, false);                                                               \
    } while (0)
#define STP_STAGED_LAUNCH(WW, G, D) hipLaunchKernelGGL((normal_residual_staged_kernel<WW, G, D>), dim3(nb), dim3(256), \
        lds, st, rows, rt, ntiles, rp, l16, val, tu, uniq, y, g, x, lam, dinv, s, z, part_sz, part_ss, done, rev)
    switch (W) {
        case 64: STP_STAGED(64); break;
        case 32: STP_STAGED(32); break;
        case 16: STP_STAGED(16); break;
        default: STP_STAGED(8); break;
    }
#undef STP_STAGED_LAUNCH
#undef STP_STAGED
    return hipGetLastError();
}

// ------------------------------------------------------------------ reductions / scalars
// Sum `n` partials into *out with one 256-thread block (fixed order: deterministic).
// Sums of up to three partial arrays (b, c optional), every thread of the
// block receives them; each thread adds its strided share in index order.
__device__ void block_sums(const double* a, int na, const double* b, int nb, const double* c, int nc, double* ra,
                           double* rb, double* rc) {
    __shared__ double red[3][4];
    double sa = 0.0, sb = 0.0, sc = 0.0;
    // Thread t adds a[t], a[t + 256], ... in that order (likewise b and c).
    // The loads of U strided entries of all three arrays go out before the
    // adds, so a thread waits for one memory latency per 256 U partials
    // instead of one per entry; the order of the adds is unchanged.
    constexpr int U = 8;
    const int nmax = na > nb ? (na > nc ? na : nc) : (nb > nc ? nb : nc);
    for (int i0 = threadIdx.x; i0 < nmax; i0 += U * 256) {
        double va[U], vb[U], vc[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = i0 + u * 256;
            va[u] = i < na ? a[i] : 0.0;
            vb[u] = b && i < nb ? b[i] : 0.0;
            vc[u] = c && i < nc ? c[i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = i0 + u * 256;
            if (i < na) sa += va[u];
            if (b && i < nb) sb += vb[u];
            if (c && i < nc) sc += vc[u];
        }
    }
    sa = wave_sum(sa);
    sb = wave_sum(sb);
    sc = wave_sum(sc);
    if ((threadIdx.x & 63) == 0) {
        const int w = threadIdx.x >> 6;
        red[0][w] = sa;
        red[1][w] = sb;
        red[2][w] = sc;
    }
    __syncthreads();
    *ra = ((red[0][0] + red[0][1]) + red[0][2]) + red[0][3];
    if (rb) *rb = ((red[1][0] + red[1][1]) + red[1][2]) + red[1][3];
    if (rc) *rc = ((red[2][0] + red[2][1]) + red[2][2]) + red[2][3];
    __syncthreads();   // red is reused by the next call
}

__device__ double block_sum(const double* part, int n) {
    double r;
    block_sums(part, n, nullptr, 0, nullptr, 0, &r, nullptr, nullptr);
    return r;
}

// x += a p (cols), r -= a q and, with HY, y += a q (rows); v32 receives the
// FP32 copy of y (HY) or r; |r|^2 and |y|^2 accumulate in rr and yy.  HC / HR:
// x and p / r and q have at least one element.  A thread's U grid-strided
// elements are loaded together (slots past an array's end load its last
// element and leave it unused), then updated in ascending order, so the sums
// keep the order of a plain strided loop.
template <bool HC, bool HR, bool HY>
__device__ __forceinline__ void cgls_update_loop(double a, int cols, int rows, double* __restrict__ x,
                                                 const float* __restrict__ p, double* __restrict__ r,
                                                 const float* __restrict__ q, double* __restrict__ y,
                                                 float* __restrict__ v32, double& rr, double& yy) {
    constexpr int U = 4;
    const int n = cols > rows ? cols : rows;
    const int stride = gridDim.x * blockDim.x;
    for (int i0 = blockIdx.x * blockDim.x + threadIdx.x; i0 < n; i0 += U * stride) {
        double xv[U], rv[U], yv[U];
        float pv[U], qv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = i0 + u * stride;
            if constexpr (HC) {
                const int ic = i < cols ? i : cols - 1;
                xv[u] = x[ic];
                pv[u] = p[ic];
            }
            if constexpr (HR) {
                const int ir = i < rows ? i : rows - 1;
                rv[u] = r[ir];
                qv[u] = q[ir];
                if constexpr (HY) yv[u] = y[ir];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = i0 + u * stride;
            if constexpr (HC)
                if (i < cols) x[i] = xv[u] + a * (double)pv[u];
            if constexpr (HR)
                if (i < rows) {
                    const double rn = rv[u] - a * (double)qv[u];
                    r[i] = rn;
                    if constexpr (HY) {
                        const double yn = yv[u] + a * (double)qv[u];
                        y[i] = yn;
                        if (v32) v32[i] = (float)yn;
                        yy += yn * yn;
                    } else if (v32) {
                        v32[i] = (float)rn;
                    }
                    rr += rn * rn;
                }
        }
    }
}

// CGLS, after q = A p (with its |q|^2 partials): alpha = gamma / (|q|^2 +
// lam |p|^2), x += alpha p (cols), r -= alpha q (rows) and, when y is given,
// y += alpha q (y = r_ref - r); v32 (when given) receives the FP32 copy of y
// (y given) or of r, and part_rr / part_yy (when given) the partials of |r|^2
// and |y|^2.  Every block sums the
// same partials in the same order, so all blocks hold the same alpha and no
// single-block scalar launch sits between the SpMV and the update; block 0
// records alpha (and the breakdown flag).  The done flag is read once per
// block (block 0 may set it while other blocks start).
__global__ void __launch_bounds__(256) cgls_alpha_update_kernel(SolverScalars* sc, const double* pq, int nq,
                                                                const double* pp, int np, double lam, int it,
                                                                int cols, double* __restrict__ x,
                                                                const float* __restrict__ p, int rows,
                                                                double* __restrict__ r, const float* __restrict__ q,
                                                                double* __restrict__ y, float* __restrict__ v32,
                                                                double* __restrict__ part_rr,
                                                                double* __restrict__ part_yy) {
    __shared__ int s_done;
    if (threadIdx.x == 0) s_done = sc->done;
    __syncthreads();
    if (s_done) return;
    double qq, ppv;
    block_sums(pq, nq, pp, np, nullptr, 0, &qq, &ppv, nullptr);
    const double del = qq + lam * ppv;
    const bool first = blockIdx.x == 0 && threadIdx.x == 0;
    if (!(del > 0.0)) {
        if (first) { sc->done = 2; sc->alpha = 0.0; }
        return;
    }
    const double a = sc->gam[it & 1] / del;
    if (first) sc->alpha = a;
    double rr = 0.0, yy = 0.0;
    if (cols > 0 && rows > 0) {
        if (y) cgls_update_loop<true, true, true>(a, cols, rows, x, p, r, q, y, v32, rr, yy);
        else cgls_update_loop<true, true, false>(a, cols, rows, x, p, r, q, y, v32, rr, yy);
    } else if (cols > 0) {
        cgls_update_loop<true, false, false>(a, cols, rows, x, p, r, q, y, v32, rr, yy);
    } else if (rows > 0) {
        if (y) cgls_update_loop<false, true, true>(a, cols, rows, x, p, r, q, y, v32, rr, yy);
        else cgls_update_loop<false, true, false>(a, cols, rows, x, p, r, q, y, v32, rr, yy);
    }
    if (part_rr) {
        block_partial(rr, part_rr);
        block_partial(yy, part_yy);
    }
}

// CGLS, after s = A^T r - lam x and z = M s (with the s.z, s.s and flexible
// partials): rel = |s| / s0, the convergence test, beta = z.s / gamma (or
// z.(s - s_prev) / gamma, flexible), then p = z + beta p with its |p|^2
// partials and s_prev <- s.  gamma sits in two slots by iteration parity:
// every block reads gam[it & 1] and block 0 writes gam[(it + 1) & 1], so no
// block reads what another block of the launch writes.  FIRST (a start or a
// restart): beta = 0, no scalar is read or written.
template <bool FIRST>
__global__ void __launch_bounds__(256) cgls_beta_direction_kernel(SolverScalars* sc, const double* psz, int nsz,
                                                                  const double* pss, int nss, const double* pflex,
                                                                  int nfl, double tol, double* history, int it,
                                                                  int cols, const float* __restrict__ z,
                                                                  float* __restrict__ p, double* __restrict__ part_pp,
                                                                  const float* __restrict__ s,
                                                                  float* __restrict__ s_prev) {
    __shared__ int s_done;
    if (threadIdx.x == 0) s_done = sc->done;
    __syncthreads();
    if (s_done) return;
    double b = 0.0;
    if constexpr (!FIRST) {
        double sz, ss, fl = 0.0;
        block_sums(psz, nsz, pss, nss, pflex, nfl, &sz, &ss, pflex ? &fl : nullptr);
        const double rel = sc->s0 > 0.0 ? sqrt(ss) / sc->s0 : 0.0;
        const bool first = blockIdx.x == 0 && threadIdx.x == 0;
        if (first) {
            sc->rel = rel;
            sc->iters = it + 1;
            if (history) history[it] = rel;
        }
        if (rel < tol) {
            if (first) sc->done = 1;
            return;
        }
        b = (pflex ? fl : sz) / sc->gam[it & 1];
        if (first) {
            sc->beta = b;
            sc->gam[(it + 1) & 1] = sz;
        }
    }
    double v = 0.0;
    constexpr int U = 4;   // loads of a thread's U strided elements together, as in the update kernel
    const int stride = gridDim.x * blockDim.x;
    for (int i0 = blockIdx.x * blockDim.x + threadIdx.x; i0 < cols; i0 += U * stride) {
        float zv[U], pv[U], sv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {   // past the end: the last element, not used
            const int ic = min(i0 + u * stride, cols - 1);
            zv[u] = z[ic];
            pv[u] = p[ic];
            if (s_prev) sv[u] = s[ic];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = i0 + u * stride;
            if (i < cols) {
                const float pn = (float)((double)zv[u] + b * (double)pv[u]);
                p[i] = pn;
                v += (double)pn * (double)pn;
                if (s_prev) s_prev[i] = sv[u];
            }
        }
    }
    block_partial(v, part_pp);
}

// partials of z.s and z.(s - s_prev) (flexible variant)
__global__ void flex_dots_kernel(const int* done, int n, const float* z, const float* s, const float* s_prev,
                                 double* part_zs, double* part_fl) {
    if (done && *done) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double a = 0.0, b = 0.0;
    if (i < n) {
        a = (double)z[i] * (double)s[i];
        b = (double)z[i] * ((double)s[i] - (double)s_prev[i]);
    }
    block_partial(a, part_zs);
    block_partial(b, part_fl);
}

// gamma <- sum(psz), s0 <- sqrt(sum(pss)) (initialisation) or only gamma
// (restart from a recomputed residual: the done flag of the finished pass is
// cleared here, so the direction kernel that follows runs)
__global__ void cgls_init_kernel(SolverScalars* sc, const double* psz, int nsz, const double* pss, int nss,
                                 int set_s0, int it) {
    double sz, ss;
    block_sums(psz, nsz, pss, nss, nullptr, 0, &sz, &ss, nullptr);
    if (threadIdx.x == 0) {
        sc->gamma = sz;
        sc->gam[it & 1] = sz;
        if (set_s0) { sc->s0 = sqrt(ss); sc->done = sc->s0 > 0.0 ? 0 : 1; sc->iters = 0; }
        else sc->done = 0;
        sc->rel = sc->s0 > 0.0 ? sqrt(ss) / sc->s0 : 0.0;
    }
}

__global__ void zero_words_kernel(uint32_t* p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0u;
}

__global__ void copy_f32_kernel(int n, const float* a, float* b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

// r = b - A x  (FP64 x, FP64 accumulation)
template <int W, typename VT>
__global__ void __launch_bounds__(256) residual_kernel(int rows, const int* __restrict__ rp, const int* __restrict__ ci,
                                                       const VT* __restrict__ val, const double* __restrict__ x,
                                                       const double* __restrict__ b, double* __restrict__ r) {
    constexpr int RPB = 256 / W;
    const int lane = threadIdx.x % W, sub = threadIdx.x / W;
    for (int row0 = blockIdx.x * RPB; row0 < rows; row0 += gridDim.x * RPB) {
        const int row = row0 + sub;
        if (row < rows) {
            const double acc = row_dot<W, double, double, VT>(rp, ci, val, x, row, lane);
            if (lane == 0) r[row] = b[row] - acc;
        }
    }
}

template <typename VT>
static hipError_t residual_t(int W, int rows, const int* rp, const int* ci, const VT* val, const double* x,
                             const double* b, double* r, hipStream_t st) {
    const int nb = spmv_blocks(rows, W);
    if (rows == 0) return hipSuccess;
    switch (W) {
        case 64: hipLaunchKernelGGL((residual_kernel<64, VT>), dim3(nb), dim3(256), 0, st, rows, rp, ci, val, x, b, r);
            break;
        case 32: hipLaunchKernelGGL((residual_kernel<32, VT>), dim3(nb), dim3(256), 0, st, rows, rp, ci, val, x, b, r);
            break;
        case 16: hipLaunchKernelGGL((residual_kernel<16, VT>), dim3(nb), dim3(256), 0, st, rows, rp, ci, val, x, b, r);
            break;
        default: hipLaunchKernelGGL((residual_kernel<8, VT>), dim3(nb), dim3(256), 0, st, rows, rp, ci, val, x, b, r);
            break;
    }
    return hipGetLastError();
}

hipError_t residual(int W, int rows, const int* rp, const int* ci, const float* val, const double* x, const double* b,
                    double* r, hipStream_t st) {
    return residual_t<float>(W, rows, rp, ci, val, x, b, r, st);
}

hipError_t residual_f64v(int W, int rows, const int* rp, const int* ci, const double* val, const double* x,
                         const double* b, double* r, hipStream_t st) {
    return residual_t<double>(W, rows, rp, ci, val, x, b, r, st);
}

// Jacobi for A^T A + lam I: dinv_j = 1 / (sum_i a_ij^2 + lam), from the rows of A^T
// (16 lanes per row of A^T, lane-strided sums combined by shuffles)
__global__ void __launch_bounds__(256) normal_diag_kernel(int n, const int* rp, const float* val, double lam,
                                                           float* dinv) {
    constexpr int W = 16;
    const int lane = threadIdx.x % W;
    const int j = blockIdx.x * (256 / W) + (int)threadIdx.x / W;
    double s = 0.0;
    if (j < n)
        for (int k = rp[j] + lane; k < rp[j + 1]; k += W) s += (double)val[k] * (double)val[k];
#pragma unroll
    for (int m = W / 2; m > 0; m >>= 1) s += __shfl_xor(s, m);
    if (j < n && lane == 0) {
        s += lam;
        dinv[j] = s > 0.0 ? (float)(1.0 / s) : 1.0f;
    }
}

// Jacobi for a square A (reference jacobi_preconditioned_cg): first diagonal
// entry of each row with |a| > 1e-30, else 1
template <typename VT>
__global__ void diag_inv_kernel(int n, const int* rp, const int* ci, const VT* val, double* dinv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double d = 1.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k)
        if (ci[k] == i) {
            if (fabs((double)val[k]) > 1e-30) d = 1.0 / (double)val[k];
            break;
        }
    dinv[i] = d;
}

// ------------------------------------------------------------------ square CG (reference semantics)
// pap = p.Ap partials are produced by spmv with a custom epilogue: reuse dot kernel
__global__ void dot_partial_kernel(const int* done, int n, const double* a, const double* b, double* part) {
    if (done && *done) return;
    double v = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) v += a[i] * b[i];
    block_partial(v, part);
}

__global__ void cg_alpha_kernel(SolverScalars* sc, const double* part, int n) {
    if (sc->done) return;
    const double pap = block_sum(part, n);
    if (threadIdx.x == 0) {
        if (fabs(pap) < 1e-30) { sc->done = 2; return; }
        sc->alpha = sc->gamma / pap;
    }
}

// x += a p; r -= a Ap; z = dinv r (or r); partials r.r and r.z
__global__ void cg_update_kernel(const SolverScalars* sc, int n, double* x, double* r, const double* p,
                                 const double* ap, const double* dinv, double* z, double* part_rr, double* part_rz) {
    if (sc->done) return;
    const double a = sc->alpha;
    double rr = 0.0, rz = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        x[i] += a * p[i];
        const double ri = r[i] - a * ap[i];
        r[i] = ri;
        const double zi = dinv ? dinv[i] * ri : ri;
        if (dinv) z[i] = zi;
        rr += ri * ri;
        rz += ri * zi;
    }
    block_partial(rr, part_rr);
    block_partial(rz, part_rz);
}

// convergence test on |r|/|b| (reference: CG and PCG both test before the p update)
__global__ void cg_beta_kernel(SolverScalars* sc, const double* prr, const double* prz, int n, double tol, int it,
                               double* history) {
    if (sc->done) return;
    double rr, rz;
    block_sums(prr, n, prz, n, nullptr, 0, &rr, &rz, nullptr);
    if (threadIdx.x == 0) {
        sc->iters = it + 1;
        sc->rel = sqrt(rr) / sc->s0;
        if (history) history[it] = sc->rel;
        if (sqrt(rr) / sc->s0 < tol) { sc->done = 1; return; }
        sc->beta = rz / sc->gamma;
        sc->gamma = rz;
    }
}

__global__ void cg_direction_kernel(const SolverScalars* sc, int n, const double* z, double* p) {
    if (sc->done) return;
    const double b = sc->beta;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = z[i] + b * p[i];
}

__global__ void cg_init_kernel(SolverScalars* sc, const double* prr, const double* prz, int n) {
    double rr, rz;
    block_sums(prr, n, prz, n, nullptr, 0, &rr, &rz, nullptr);
    if (threadIdx.x == 0) {
        sc->s0 = sqrt(rr);
        sc->gamma = rz;
        sc->iters = 0;
        sc->rel = sc->s0 < 1e-30 ? 0.0 : 1.0;
        sc->done = sc->s0 < 1e-30 ? 1 : 0;
    }
}

// ------------------------------------------------------------------ Siddon rays
#pragma clang fp contract(off)
struct RayWalk {
    double p0[3], d[3], len, tnext[3];
    int64_t knext[3], kend[3], kdir[3];
    double h, t0;
    int64_t n[3];
    bool fin;
};

__device__ __forceinline__ bool axis_exhausted(int64_t k, int64_t kend, int64_t kdir) {
    return kdir > 0 ? k > kend : k < kend;
}

__device__ void ray_init(RayWalk& w, const double* pt, const int64_t* n, double h) {
    double l2 = 0.0;
    for (int a = 0; a < 3; ++a) {
        w.p0[a] = pt[a];
        w.d[a] = pt[3 + a] - pt[a];
        l2 += w.d[a] * w.d[a];
        w.n[a] = n[a];
    }
    w.len = sqrt(l2);
    w.h = h;
    w.t0 = 0.0;
    w.fin = w.len == 0.0;
    for (int a = 0; a < 3; ++a) {
        const double p1 = pt[3 + a];
        if (w.d[a] > 0) {
            w.kdir[a] = 1;
            w.knext[a] = (int64_t)floor(w.p0[a] / h) + 1;
            w.kend[a] = (int64_t)ceil(p1 / h) - 1;
        } else if (w.d[a] < 0) {
            w.kdir[a] = -1;
            w.knext[a] = (int64_t)ceil(w.p0[a] / h) - 1;
            w.kend[a] = (int64_t)floor(p1 / h) + 1;
        } else {
            w.kdir[a] = 0; w.knext[a] = 0; w.kend[a] = 0;
        }
        w.tnext[a] = w.d[a] != 0 ? ((double)w.knext[a] * h - w.p0[a]) / w.d[a] : INFINITY;
        if (w.d[a] != 0 && axis_exhausted(w.knext[a], w.kend[a], w.kdir[a])) w.tnext[a] = INFINITY;
    }
}

// next segment; returns false when the walk is over
__device__ bool ray_next(RayWalk& w, int64_t* cell, double* seg) {
    while (!w.fin) {
        int a = 0;
        if (w.tnext[1] < w.tnext[a]) a = 1;
        if (w.tnext[2] < w.tnext[a]) a = 2;
        const double t1 = w.tnext[a] < 1.0 ? w.tnext[a] : 1.0;
        const double t0 = w.t0;
        bool emit = false;
        if (t1 > t0) {
            const double tm = 0.5 * (t0 + t1);
            int64_t ix = (int64_t)floor((w.p0[0] + tm * w.d[0]) / w.h);
            int64_t iy = (int64_t)floor((w.p0[1] + tm * w.d[1]) / w.h);
            int64_t iz = (int64_t)floor((w.p0[2] + tm * w.d[2]) / w.h);
            ix = ix < 0 ? 0 : (ix >= w.n[0] ? w.n[0] - 1 : ix);
            iy = iy < 0 ? 0 : (iy >= w.n[1] ? w.n[1] - 1 : iy);
            iz = iz < 0 ? 0 : (iz >= w.n[2] ? w.n[2] - 1 : iz);
            *cell = (ix * w.n[1] + iy) * w.n[2] + iz;
            *seg = (t1 - t0) * w.len;
            emit = true;
        }
        if (t1 >= 1.0) {
            w.fin = true;
        } else {
            w.t0 = t1;
            w.knext[a] += w.kdir[a];
            w.tnext[a] = axis_exhausted(w.knext[a], w.kend[a], w.kdir[a])
                             ? INFINITY
                             : ((double)w.knext[a] * w.h - w.p0[a]) / w.d[a];
        }
        if (emit) return true;
    }
    return false;
}
#pragma clang fp contract(on)

__global__ void rays_count_kernel(int64_t nrays, const double* pts, int64_t n0, int64_t n1, int64_t n2, double h,
                                  int* counts) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrays) return;
    const int64_t n[3] = {n0, n1, n2};
    RayWalk w;
    ray_init(w, pts + 6 * r, n, h);
    int c = 0;
    int64_t cell;
    double seg;
    while (ray_next(w, &cell, &seg)) ++c;
    counts[r] = c;
}

// fill the row in traversal order, then insertion-sort it by column
__global__ void rays_fill_kernel(int64_t nrays, const double* pts, int64_t n0, int64_t n1, int64_t n2, double h,
                                 const int* rp, int* cols, float* vals) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrays) return;
    const int64_t n[3] = {n0, n1, n2};
    RayWalk w;
    ray_init(w, pts + 6 * r, n, h);
    const int b = rp[r], e = rp[r + 1];
    int k = b;
    int64_t cell;
    double seg;
    while (k < e && ray_next(w, &cell, &seg)) {
        const int c = (int)cell;
        const float v = (float)seg;
        int j = k;
        while (j > b && cols[j - 1] > c) { cols[j] = cols[j - 1]; vals[j] = vals[j - 1]; --j; }
        cols[j] = c;
        vals[j] = v;
        ++k;
    }
}

hipError_t rays_count(int64_t nrays, const double* pts, const int64_t* n, double h, int* counts, hipStream_t st) {
    const unsigned nb = (unsigned)((nrays + 127) / 128);
    if (nb) hipLaunchKernelGGL(rays_count_kernel, dim3(nb), dim3(128), 0, st, nrays, pts, n[0], n[1], n[2], h, counts);
    return hipGetLastError();
}

hipError_t rays_fill(int64_t nrays, const double* pts, const int64_t* n, double h, const int* rp, int* cols,
                     float* vals, hipStream_t st) {
    const unsigned nb = (unsigned)((nrays + 127) / 128);
    if (nb)
        hipLaunchKernelGGL(rays_fill_kernel, dim3(nb), dim3(128), 0, st, nrays, pts, n[0], n[1], n[2], h, rp, cols,
                           vals);
    return hipGetLastError();
}

// ------------------------------------------------------------------ transpose
__global__ void t_keys_kernel(int rows, const int* rp, const int* ci, uint64_t* keys, int* counts) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    for (int k = rp[r]; k < rp[r + 1]; ++k) {
        keys[k] = ((uint64_t)(uint32_t)ci[k] << 32) | (uint32_t)r;
        atomicAdd(&counts[ci[k]], 1);
    }
}

__global__ void t_cols_kernel(int64_t nnz, const uint64_t* keys, int* tci) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nnz) tci[i] = (int)(keys[i] & 0xffffffffu);
}

hipError_t exclusive_scan_i32(const int* in, int* out, int n, hipStream_t st) {
    size_t tb = 0;
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, tb, in, out, n, st);
    if (e != hipSuccess) return e;
    void* tmp = nullptr;
    if ((e = hipMalloc(&tmp, tb ? tb : 1)) != hipSuccess) return e;
    e = hipcub::DeviceScan::ExclusiveSum(tmp, tb, in, out, n, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(tmp);
    return e;
}

// A (rows x cols) -> A^T (cols x rows), entries of a column ordered by row
hipError_t csr_transpose(int rows, int cols, int64_t nnz, const int* rp, const int* ci, const float* val, int* trp,
                         int* tci, float* tval, hipStream_t st) {
    hipError_t e = hipSuccess;
    uint64_t *k0 = nullptr, *k1 = nullptr;
    float* v1 = nullptr;
    int* counts = nullptr;
    void* tmp = nullptr;
    size_t tb = 0;
    do {
        if ((e = hipMalloc(&counts, sizeof(int) * ((size_t)cols + 1))) != hipSuccess) break;
        if ((e = hipMemsetAsync(counts, 0, sizeof(int) * ((size_t)cols + 1), st)) != hipSuccess) break;
        if (nnz > 0) {
            if ((e = hipMalloc(&k0, sizeof(uint64_t) * nnz)) != hipSuccess) break;
            if ((e = hipMalloc(&k1, sizeof(uint64_t) * nnz)) != hipSuccess) break;
            if ((e = hipMalloc(&v1, sizeof(float) * nnz)) != hipSuccess) break;
            hipLaunchKernelGGL(t_keys_kernel, dim3((rows + 255) / 256), dim3(256), 0, st, rows, rp, ci, k0, counts);
            if ((e = hipGetLastError()) != hipSuccess) break;
            if ((e = hipcub::DeviceRadixSort::SortPairs(nullptr, tb, k0, k1, val, v1, (int)nnz, 0, 64, st)) != hipSuccess)
                break;
            if ((e = hipMalloc(&tmp, tb ? tb : 1)) != hipSuccess) break;
            if ((e = hipcub::DeviceRadixSort::SortPairs(tmp, tb, k0, k1, val, v1, (int)nnz, 0, 64, st)) != hipSuccess)
                break;
            hipLaunchKernelGGL(t_cols_kernel, dim3((unsigned)((nnz + 255) / 256)), dim3(256), 0, st, nnz, k1, tci);
            if ((e = hipGetLastError()) != hipSuccess) break;
            if ((e = hipMemcpyAsync(tval, v1, sizeof(float) * nnz, hipMemcpyDeviceToDevice, st)) != hipSuccess) break;
        }
        if ((e = exclusive_scan_i32(counts, trp, cols + 1, st)) != hipSuccess) break;
    } while (0);
    {   // report a failure of the queued work unless an earlier one is reported
        const hipError_t es = hipStreamSynchronize(st);
        if (e == hipSuccess) e = es;
    }
    (void)hipFree(tmp); (void)hipFree(k0); (void)hipFree(k1); (void)hipFree(v1); (void)hipFree(counts);
    return e;
}

// ------------------------------------------------------------------ launch helpers
static unsigned nb256(int64_t n) { return (unsigned)((n + 255) / 256); }

int cgls_blocks(int n) {
    int b = (n + 255) / 256;
    return b > 256 ? 256 : (b < 1 ? 1 : b);
}
hipError_t launch_cgls_alpha_update(SolverScalars* sc, const double* pq, int nq, const double* pp, int np,
                                    double lam, int it, int cols, double* x, const float* p, int rows, double* r,
                                    const float* q, double* y, float* v32, double* part_rr, double* part_yy,
                                    hipStream_t st) {
    hipLaunchKernelGGL(cgls_alpha_update_kernel, dim3(cgls_blocks(cols > rows ? cols : rows)), dim3(256), 0, st, sc,
                       pq, nq, pp, np, lam, it, cols, x, p, rows, r, q, y, v32, part_rr, part_yy);
    return hipGetLastError();
}
hipError_t launch_cgls_beta_direction(SolverScalars* sc, const double* psz, int nsz, const double* pss, int nss,
                                      const double* pflex, int nfl, double tol, double* history, int it, int first,
                                      int cols, const float* z, float* p, double* part_pp, const float* s,
                                      float* s_prev, hipStream_t st) {
    if (first)
        hipLaunchKernelGGL(cgls_beta_direction_kernel<true>, dim3(cgls_blocks(cols)), dim3(256), 0, st, sc, psz, nsz,
                           pss, nss, pflex, nfl, tol, history, it, cols, z, p, part_pp, s, s_prev);
    else
        hipLaunchKernelGGL(cgls_beta_direction_kernel<false>, dim3(cgls_blocks(cols)), dim3(256), 0, st, sc, psz,
                           nsz, pss, nss, pflex, nfl, tol, history, it, cols, z, p, part_pp, s, s_prev);
    return hipGetLastError();
}
hipError_t launch_flex_dots(const int* done, int n, const float* z, const float* s, const float* s_prev,
                            double* part_zs, double* part_fl, hipStream_t st) {
    hipLaunchKernelGGL(flex_dots_kernel, dim3(nb256(n)), dim3(256), 0, st, done, n, z, s, s_prev, part_zs, part_fl);
    return hipGetLastError();
}
hipError_t launch_cgls_init(SolverScalars* sc, const double* psz, int nsz, const double* pss, int nss, int set_s0,
                            int it, hipStream_t st) {
    hipLaunchKernelGGL(cgls_init_kernel, dim3(1), dim3(256), 0, st, sc, psz, nsz, pss, nss, set_s0, it);
    return hipGetLastError();
}
hipError_t launch_zero_words(void* p, size_t nwords, hipStream_t st) {
    if (nwords == 0) return hipSuccess;
    size_t nb = (nwords + 255) / 256;
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL(zero_words_kernel, dim3((unsigned)nb), dim3(256), 0, st, static_cast<uint32_t*>(p), nwords);
    return hipGetLastError();
}
hipError_t launch_copy_f32(int n, const float* a, float* b, hipStream_t st) {
    hipLaunchKernelGGL(copy_f32_kernel, dim3(nb256(n)), dim3(256), 0, st, n, a, b);
    return hipGetLastError();
}
hipError_t launch_normal_diag(int n, const int* rp, const float* val, double lam, float* dinv, hipStream_t st) {
    hipLaunchKernelGGL(normal_diag_kernel, dim3((unsigned)((n + 15) / 16 > 0 ? (n + 15) / 16 : 1)), dim3(256), 0, st,
                       n, rp, val, lam, dinv);
    return hipGetLastError();
}
hipError_t launch_diag_inv(int n, const int* rp, const int* ci, const float* val, double* dinv, hipStream_t st) {
    hipLaunchKernelGGL(diag_inv_kernel<float>, dim3(nb256(n)), dim3(256), 0, st, n, rp, ci, val, dinv);
    return hipGetLastError();
}

hipError_t launch_diag_inv_f64v(int n, const int* rp, const int* ci, const double* val, double* dinv,
                                hipStream_t st) {
    hipLaunchKernelGGL(diag_inv_kernel<double>, dim3(nb256(n)), dim3(256), 0, st, n, rp, ci, val, dinv);
    return hipGetLastError();
}
int cg_blocks(int n) {
    int b = (n + 255) / 256;
    return b > 1024 ? 1024 : (b < 1 ? 1 : b);
}
hipError_t launch_dot_partial(const int* done, int n, const double* a, const double* b, double* part, hipStream_t st) {
    hipLaunchKernelGGL(dot_partial_kernel, dim3(cg_blocks(n)), dim3(256), 0, st, done, n, a, b, part);
    return hipGetLastError();
}
hipError_t launch_cg_alpha(SolverScalars* sc, const double* part, int n, hipStream_t st) {
    hipLaunchKernelGGL(cg_alpha_kernel, dim3(1), dim3(256), 0, st, sc, part, n);
    return hipGetLastError();
}
hipError_t launch_cg_update(const SolverScalars* sc, int n, double* x, double* r, const double* p, const double* ap,
                            const double* dinv, double* z, double* prr, double* prz, hipStream_t st) {
    hipLaunchKernelGGL(cg_update_kernel, dim3(cg_blocks(n)), dim3(256), 0, st, sc, n, x, r, p, ap, dinv, z, prr, prz);
    return hipGetLastError();
}
hipError_t launch_cg_beta(SolverScalars* sc, const double* prr, const double* prz, int n, double tol, int it,
                          double* history, hipStream_t st) {
    hipLaunchKernelGGL(cg_beta_kernel, dim3(1), dim3(256), 0, st, sc, prr, prz, n, tol, it, history);
    return hipGetLastError();
}
hipError_t launch_cg_direction(const SolverScalars* sc, int n, const double* z, double* p, hipStream_t st) {
    hipLaunchKernelGGL(cg_direction_kernel, dim3(cg_blocks(n)), dim3(256), 0, st, sc, n, z, p);
    return hipGetLastError();
}
hipError_t launch_cg_init(SolverScalars* sc, const double* prr, const double* prz, int n, hipStream_t st) {
    hipLaunchKernelGGL(cg_init_kernel, dim3(1), dim3(256), 0, st, sc, prr, prz, n);
    return hipGetLastError();
}

}  // namespace stp
