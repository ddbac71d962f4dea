#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace stp {

// Device-resident solver scalars (one struct per solve).
struct SolverScalars {
    double alpha, beta, gamma, s0, rel;
    int done;    // 0 running, 1 converged, 2 breakdown
    int iters;
    double gam[2];   // CGLS: gamma of iteration it in gam[it & 1]
};

// upper bound of the grid (and so of the per-block partial sums) of the SpMV
// and normal-residual kernels: they stride over the rows
constexpr int SPMV_MAX_BLOCKS = 2048;
// entries of the LDS staging buffer of the staged A^T product (FP32: 32 KB)
constexpr int STAGE_CAP = 8192;

int spmv_lanes(double mean_nnz, int transpose);
int spmv_lanes_layout(double mean_nnz, int transpose);
int spmv_blocks(int rows, int W);
// rev != 0: rows visited last to first (same per-row results)
hipError_t spmv_f32(int W, int rows, const int* rp, const int* ci, const float* val, const float* x, float* y,
                    double* part, const int* done, hipStream_t st, int rev = 0);
hipError_t spmv_f64(int W, int rows, const int* rp, const int* ci, const float* val, const double* x, double* y,
                    double alpha, const double* wv, double lam, double* part, const int* done, hipStream_t st);
// s = A^T r - lam x (FP64 r); gout, when given, receives A^T r in FP64
hipError_t normal_residual(int W, int rows, const int* rp, const int* ci, const float* val, const double* r,
                           const double* x, double lam, const float* dinv, float* s, float* z, double* part_sz,
                           double* part_ss, const int* done, hipStream_t st, int rev = 0, double* gout = nullptr);
// s = g - A^T y - lam x with the change y = r_ref - r of the residual gathered
// in FP32 and g = A^T r_ref in FP64; without g, s = A^T y - lam x with y = r
// in FP32 (FP64 accumulation either way)
hipError_t normal_residual_r32(int W, int rows, const int* rp, const int* ci, const float* val, const float* y,
                               const double* g, const double* x, double lam, const float* dinv, float* s, float* z,
                               double* part_sz, double* part_ss, const int* done, int rev, hipStream_t st);
// the same from the staged A^T (tiles of rt rows; per tile its distinct
// entries of y are gathered once into LDS); the grid is
// staged_blocks(ntiles), which is also the count of partials
int staged_blocks(int ntiles);
hipError_t normal_residual_staged(int W, int rows, int rt, int ntiles, int maxu, const int* rp, const uint16_t* l16,
                                  const float* val, const int* tu, const int* uniq, const float* y, const double* g,
                                  const double* x, double lam, const float* dinv, float* s, float* z,
                                  double* part_sz, double* part_ss, const int* done, int rev, hipStream_t st);
hipError_t residual(int W, int rows, const int* rp, const int* ci, const float* val, const double* x, const double* b,
                    double* r, hipStream_t st);
// FP64-valued variants (the square CG drop-in when FP64 values are attached)
hipError_t spmv_f64v(int W, int rows, const int* rp, const int* ci, const double* val, const double* x, double* y,
                     const int* done, hipStream_t st);
hipError_t residual_f64v(int W, int rows, const int* rp, const int* ci, const double* val, const double* x,
                         const double* b, double* r, hipStream_t st);
// the SpMV without its value stream (measurement of the gather cost)
template <typename XT>
hipError_t gather_probe(int W, int rows, const int* rp, const int* ci, const XT* x, XT* y, hipStream_t st);

hipError_t rays_count(int64_t nrays, const double* pts, const int64_t* n, double h, int* counts, hipStream_t st);
hipError_t rays_fill(int64_t nrays, const double* pts, const int64_t* n, double h, const int* rp, int* cols,
                     float* vals, hipStream_t st);
hipError_t exclusive_scan_i32(const int* in, int* out, int n, hipStream_t st);
hipError_t csr_transpose(int rows, int cols, int64_t nnz, const int* rp, const int* ci, const float* val, int* trp,
                         int* tci, float* tval, hipStream_t st);

// CGLS iteration kernels.  cgls_blocks(n) bounds the grid of the
// elementwise CGLS kernels (grid-stride), and so the count of |p|^2 partials.
int cgls_blocks(int n);
// alpha = gam[it&1] / (|q|^2 + lam |p|^2) (every block sums the partials),
// x += alpha p, r -= alpha q, and y += alpha q when y is given; v32 (when
// given) receives the FP32 copy of y (y given) or of r; part_rr / part_yy
// (when given, cgls_blocks(max(rows, cols)) each) the partials of |r|^2, |y|^2
hipError_t launch_cgls_alpha_update(SolverScalars* sc, const double* pq, int nq, const double* pp, int np,
                                    double lam, int it, int cols, double* x, const float* p, int rows, double* r,
                                    const float* q, double* y, float* v32, double* part_rr, double* part_yy,
                                    hipStream_t st);
// rel, convergence test, beta = (z.s or z.(s - s_prev)) / gam[it&1] (every
// block), gam[(it+1)&1] = z.s, then p = z + beta p with |p|^2 partials and
// s_prev <- s; with first = 1 (a start or restart) beta = 0 and no scalar is
// read or written
hipError_t launch_cgls_beta_direction(SolverScalars* sc, const double* psz, int nsz, const double* pss, int nss,
                                      const double* pflex, int nfl, double tol, double* history, int it, int first,
                                      int cols, const float* z, float* p, double* part_pp, const float* s,
                                      float* s_prev, hipStream_t st);
hipError_t launch_flex_dots(const int* done, int n, const float* z, const float* s, const float* s_prev,
                            double* part_zs, double* part_fl, hipStream_t st);
hipError_t launch_cgls_init(SolverScalars* sc, const double* psz, int nsz, const double* pss, int nss, int set_s0,
                            int it, hipStream_t st);
hipError_t launch_copy_f32(int n, const float* a, float* b, hipStream_t st);
// zero n 32-bit words with a kernel (no memset node when a caller captures)
hipError_t launch_zero_words(void* p, size_t nwords, hipStream_t st);
hipError_t launch_normal_diag(int n, const int* rp, const float* val, double lam, float* dinv, hipStream_t st);
hipError_t launch_diag_inv(int n, const int* rp, const int* ci, const float* val, double* dinv, hipStream_t st);
hipError_t launch_diag_inv_f64v(int n, const int* rp, const int* ci, const double* val, double* dinv,
                                hipStream_t st);
int cg_blocks(int n);
hipError_t launch_dot_partial(const int* done, int n, const double* a, const double* b, double* part, hipStream_t st);
hipError_t launch_cg_alpha(SolverScalars* sc, const double* part, int n, hipStream_t st);
hipError_t launch_cg_update(const SolverScalars* sc, int n, double* x, double* r, const double* p, const double* ap,
                            const double* dinv, double* z, double* prr, double* prz, hipStream_t st);
hipError_t launch_cg_beta(SolverScalars* sc, const double* prr, const double* prz, int n, double tol, int it,
                          double* history, hipStream_t st);
hipError_t launch_cg_direction(const SolverScalars* sc, int n, const double* z, double* p, hipStream_t st);
hipError_t launch_cg_init(SolverScalars* sc, const double* prr, const double* prz, int n, hipStream_t st);

}  // namespace stp
