// Sparse solvers (drop-in for the reference's src/cpp/sparse/sparse_ops.h:9-48).
// CG and Jacobi-PCG keep the reference's signatures and stopping rules and
// run on the GPU (stp_cg: FP64 vectors, dots and matrix values, x0 = 0).
// CGLS on the damped normal equations is new.  smooth_gradient,
// gaussian_filter_3d and compute_directional_derivative run on the GPU too.
#pragma once

#include <cstdint>
#include <vector>

#include "csr_matrix.h"

namespace seismic {

struct SolveInfo {
    int iterations = 0;
    bool converged = false;
    double rel_residual = 0.0;        // recursive residual at exit
    double true_rel_residual = 0.0;   // recomputed in FP64 at exit
    int replacements = 0;             // FP64 residual replacements (CGLS)
};

enum class CglsPreconditioner { NONE = 0, JACOBI = 1, LEARNED = 2 };

class SparseOps {
public:
    // stop when |r| / |b| < tolerance; breakdown when |p.Ap| < 1e-30
    // (sparse_ops.cpp:8-50)
    static AlignedVector<double> conjugate_gradient(const CSRMatrix& A, const AlignedVector<double>& b, int max_iter,
                                                    double tolerance, SolveInfo* info = nullptr);
    // diagonal = first entry with col == row and |a| > 1e-30, else 1
    // (sparse_ops.cpp:52-113)
    static AlignedVector<double> jacobi_preconditioned_cg(const CSRMatrix& A, const AlignedVector<double>& b,
                                                          int max_iter, double tolerance, SolveInfo* info = nullptr);
    // new: CGLS on (A^T A + lambda I) x = A^T b to |A^T r - lambda x| / |A^T b|
    // < tolerance; `network` is an stp_lpc* prepared for A when precond is
    // LEARNED
    static AlignedVector<double> cgls(const CSRMatrix& A, const AlignedVector<double>& b, double lambda,
                                      double tolerance, int max_iter, CglsPreconditioner precond,
                                      SolveInfo* info = nullptr, void* network = nullptr);

    // Gaussian smoothing, radius (int)(3 sigma), weights renormalised over the
    // in-range cells (sparse_ops.cpp:115-158); a gradient shorter than
    // nx*ny*nz gives zeros.  sigma > 0, radius <= 64.
    static std::vector<double> smooth_gradient(const std::vector<double>& gradient, int64_t nx, int64_t ny,
                                               int64_t nz, double sigma);

    // (A direction)[i] for i < gradient.size() (sparse_ops.cpp:160-170); the
    // reference reads past A's rows when gradient is longer, here those
    // entries are 0
    static std::vector<double> compute_directional_derivative(const CSRMatrix& A,
                                                              const AlignedVector<double>& direction,
                                                              const AlignedVector<double>& gradient);

    // the normalised Gaussian exp(-|c|^2 / (2 sigma^2)) of every cell of an
    // nx x ny x nz grid, c = index - n / 2.0 (sparse_ops.cpp:212-234)
    static std::vector<double> gaussian_filter_3d(int64_t nx, int64_t ny, int64_t nz, double sigma);

    // FP64 dot product over min(|a|, |b|) entries (sparse_ops.cpp:172-187)
    static double dot_product(const AlignedVector<double>& a, const AlignedVector<double>& b);
    static double l2_norm(const AlignedVector<double>& a);
    // alpha x + y into a new vector (sparse_ops.cpp:193-202)
    static AlignedVector<double> axpy(double alpha, const AlignedVector<double>& x, const AlignedVector<double>& y);
    // A + regularization I (sparse_ops.cpp:204-210)
    static CSRMatrix build_hessian_approx(const CSRMatrix& forward_operator, double regularization);
};

}  // namespace seismic
