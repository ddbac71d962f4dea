// SparseOps (reference interface: src/cpp/sparse/sparse_ops.h:9-48): the
// solvers and the dot product run on the GPU through the C ABI.
#include "sparse/sparse_ops.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "device.h"

namespace seismic {

namespace {

void fill_info(SolveInfo* info, const stp_solve_info& s) {
    if (!info) return;
    info->iterations = s.iterations;
    info->converged = s.converged != 0;
    info->rel_residual = s.rel_residual;
    info->true_rel_residual = s.true_rel_residual;
    info->replacements = s.replacements;
}

AlignedVector<double> cg(const CSRMatrix& A, const AlignedVector<double>& b, int max_iter, double tol, int jacobi,
                         SolveInfo* info) {
    if (A.rows() != A.cols()) throw std::invalid_argument("SparseOps CG: the matrix must be square");
    const int64_t n = A.rows();
    if (static_cast<int64_t>(b.size()) < n) throw std::invalid_argument("SparseOps CG: b is too short");
    AlignedVector<double> x(static_cast<size_t>(n), 0.0);
    if (n == 0) return x;
    dev::Buffer<double> db(n), dx(n);
    dev::upload(db.get(), b.data(), n);
    stp_solve_info s{};
    dev::check(stp_cg(static_cast<const stp_csr*>(A.device_handle()), db.get(), max_iter, tol, jacobi, dx.get(),
                      nullptr, &s, nullptr),
               "SparseOps CG");
    dev::download(x.data(), dx.get(), n);
    fill_info(info, s);
    return x;
}

}  // namespace

AlignedVector<double> SparseOps::conjugate_gradient(const CSRMatrix& A, const AlignedVector<double>& b, int max_iter,
                                                    double tolerance, SolveInfo* info) {
    return cg(A, b, max_iter, tolerance, 0, info);
}

AlignedVector<double> SparseOps::jacobi_preconditioned_cg(const CSRMatrix& A, const AlignedVector<double>& b,
                                                          int max_iter, double tolerance, SolveInfo* info) {
    return cg(A, b, max_iter, tolerance, 1, info);
}

AlignedVector<double> SparseOps::cgls(const CSRMatrix& A, const AlignedVector<double>& b, double lambda,
                                      double tolerance, int max_iter, CglsPreconditioner precond, SolveInfo* info,
                                      void* network) {
    if (static_cast<int64_t>(b.size()) < A.rows()) throw std::invalid_argument("SparseOps::cgls: b is too short");
    AlignedVector<double> x(static_cast<size_t>(A.cols()), 0.0);
    if (A.cols() == 0) return x;
    dev::Buffer<double> db(A.rows()), dx(A.cols());
    dev::upload(db.get(), b.data(), A.rows());
    stp_solve_info s{};
    dev::check(stp_cgls(static_cast<const stp_csr*>(A.device_handle()), db.get(), lambda, tolerance, max_iter,
                        static_cast<int32_t>(precond), network, dx.get(), nullptr, &s, nullptr),
               "SparseOps::cgls");
    dev::download(x.data(), dx.get(), A.cols());
    fill_info(info, s);
    return x;
}

std::vector<double> SparseOps::smooth_gradient(const std::vector<double>& gradient, int64_t nx, int64_t ny,
                                               int64_t nz, double sigma) {
    if (nx < 0 || ny < 0 || nz < 0) throw std::invalid_argument("SparseOps::smooth_gradient: negative extent");
    const int64_t n = nx * ny * nz;
    std::vector<double> out(static_cast<size_t>(n), 0.0);
    if (n == 0 || static_cast<int64_t>(gradient.size()) < n) return out;
    dev::Buffer<double> din(n), dout(n), work(n);
    dev::upload(din.get(), gradient.data(), n);
    dev::check(stp_smooth_gradient_f64(nx, ny, nz, sigma, din.get(), dout.get(), work.get(), nullptr),
               "SparseOps::smooth_gradient");
    dev::download(out.data(), dout.get(), n);
    return out;
}

std::vector<double> SparseOps::compute_directional_derivative(const CSRMatrix& A,
                                                              const AlignedVector<double>& direction,
                                                              const AlignedVector<double>& gradient) {
    std::vector<double> out(gradient.size(), 0.0);
    if (out.empty()) return out;
    const AlignedVector<double> ad = A.spmv(direction);
    const size_t m = std::min(out.size(), ad.size());
    for (size_t i = 0; i < m; ++i) out[i] = ad[i];
    return out;
}

std::vector<double> SparseOps::gaussian_filter_3d(int64_t nx, int64_t ny, int64_t nz, double sigma) {
    if (nx < 0 || ny < 0 || nz < 0) throw std::invalid_argument("SparseOps::gaussian_filter_3d: negative extent");
    const int64_t n = nx * ny * nz;
    std::vector<double> out(static_cast<size_t>(n), 0.0);
    if (n == 0) return out;
    dev::Buffer<double> dout(n), work(1024);
    dev::check(stp_gaussian_filter_3d_f64(nx, ny, nz, sigma, dout.get(), work.get(), nullptr),
               "SparseOps::gaussian_filter_3d");
    dev::download(out.data(), dout.get(), n);
    return out;
}

double SparseOps::dot_product(const AlignedVector<double>& a, const AlignedVector<double>& b) {
    const int64_t n = static_cast<int64_t>(std::min(a.size(), b.size()));
    if (n == 0) return 0.0;
    dev::Buffer<double> da(n), dbuf(n);
    dev::upload(da.get(), a.data(), n);
    dev::upload(dbuf.get(), b.data(), n);
    double out = 0.0;
    dev::check(stp_dot_f64(n, da.get(), dbuf.get(), &out, nullptr), "SparseOps::dot_product");
    return out;
}

double SparseOps::l2_norm(const AlignedVector<double>& a) { return std::sqrt(dot_product(a, a)); }

AlignedVector<double> SparseOps::axpy(double alpha, const AlignedVector<double>& x, const AlignedVector<double>& y) {
    const size_t n = std::min(x.size(), y.size());
    AlignedVector<double> out(n);
    for (size_t i = 0; i < n; ++i) out[i] = alpha * x[i] + y[i];
    return out;
}

CSRMatrix SparseOps::build_hessian_approx(const CSRMatrix& forward_operator, double regularization) {
    CSRMatrix h = forward_operator.clone();
    h.add_diagonal(regularization);
    return h;
}

}  // namespace seismic
