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
