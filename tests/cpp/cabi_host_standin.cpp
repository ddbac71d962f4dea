// CPU stand-in of the part of the C ABI (include/seismic_hip.h) that the C++
// CSRMatrix / SparseOps layer calls, so that layer and dropin_driver.cpp can
// be built for the host alone and run under AddressSanitizer with the sizes
// and values a real solve produces.  Test infrastructure only.
//
// "Device" memory is host heap registered by address.  Every pointer the
// layer passes as a device pointer must lie inside a live device allocation
// with room for the access, every pointer it passes as a host pointer must
// not be a device one, and every CSR array entering stp_csr_create is checked
// (row offsets from 0, non-decreasing, ending at nnz; columns in range).  A
// violation prints what was wrong and aborts, since the library on a GPU
// would read or write out of bounds there.  Arithmetic comes from the FP64
// oracle (oracle/sparse_oracle.c): this checks the layer's host logic and
// data flow, not the kernels.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "seismic_hip.h"

extern "C" {
int64_t orc_csr_from_triplets(int64_t rows, int64_t cols, int64_t n, const int64_t* r, const int64_t* c,
                              const double* v, int64_t* row_ptr, int64_t* col_out, double* val_out);
void orc_csr_spmv(int64_t rows, const int64_t* rp, const int64_t* ci, const double* val, const double* x,
                  double* y);
void orc_csr_spmv_t(int64_t rows, int64_t cols, const int64_t* rp, const int64_t* ci, const double* val,
                    const double* x, double* y);
double orc_dot(const double* a, const double* b, int64_t n);
int orc_cg(int64_t n, const int64_t* rp, const int64_t* ci, const double* val, const double* b, int max_iter,
           double tol, double* x, double* relres);
int orc_jacobi_pcg(int64_t n, const int64_t* rp, const int64_t* ci, const double* val, const double* b,
                   int max_iter, double tol, double* x, double* relres);
int orc_cgls(int64_t rows, int64_t cols, const int64_t* rp, const int64_t* ci, const double* val, const double* b,
             double lam, double tol, int max_iter, int precond, double* x, double* relres, double* history);
void orc_smooth_gradient(int64_t nx, int64_t ny, int64_t nz, double sigma, const double* g, double* out);
}

struct stp_csr {
    int64_t rows = 0, cols = 0, nnz = 0;
    std::vector<int64_t> rp, ci;
    std::vector<double> v32;   // the FP32 values, widened
    std::vector<double> v64;   // FP64 values when attached
    int64_t grid[3] = {0, 0, 0};
};

namespace {

std::map<uintptr_t, size_t>& live() {
    static std::map<uintptr_t, size_t> m;
    return m;
}
thread_local std::string g_err;

[[noreturn]] void die(const char* what, const void* p, int64_t bytes) {
    std::fprintf(stderr, "C ABI stand-in: %s (pointer %p, %lld bytes)\n", what, p, (long long)bytes);
    std::fflush(stderr);
    std::abort();
}

// p must point into a live device allocation with `bytes` bytes from p on
void dev(const void* p, int64_t bytes, const char* what) {
    if (bytes < 0) die(what, p, bytes);
    if (bytes == 0) return;
    auto& m = live();
    auto it = m.upper_bound(reinterpret_cast<uintptr_t>(p));
    if (it == m.begin()) die(what, p, bytes);
    --it;
    const uintptr_t a = reinterpret_cast<uintptr_t>(p);
    if (a < it->first || a + static_cast<uintptr_t>(bytes) > it->first + it->second) die(what, p, bytes);
}

// p must be a host pointer (not inside a device allocation)
void host(const void* p, const char* what) {
    if (!p) die(what, p, 0);
    auto& m = live();
    auto it = m.upper_bound(reinterpret_cast<uintptr_t>(p));
    if (it == m.begin()) return;
    --it;
    if (reinterpret_cast<uintptr_t>(p) < it->first + it->second) die(what, p, 0);
}

int fail(int code, const char* msg) {
    g_err = msg;
    return code;
}

const std::vector<double>& values(const stp_csr* A) { return A->v64.empty() ? A->v32 : A->v64; }

}  // namespace

extern "C" {

const char* stp_last_error(void) { return g_err.c_str(); }

int stp_device_alloc(void** p, int64_t bytes) {
    host(p, "stp_device_alloc: result pointer");
    if (bytes <= 0) return fail(STP_ERR_INVALID, "stp_device_alloc: size");
    void* q = std::malloc(static_cast<size_t>(bytes));
    if (!q) return fail(STP_ERR_ALLOC, "stp_device_alloc: out of memory");
    live()[reinterpret_cast<uintptr_t>(q)] = static_cast<size_t>(bytes);
    *p = q;
    return STP_OK;
}

int stp_device_free(void* p) {
    if (!p) return STP_OK;
    auto& m = live();
    auto it = m.find(reinterpret_cast<uintptr_t>(p));
    if (it == m.end()) die("stp_device_free: not the start of a device allocation", p, 0);
    m.erase(it);
    std::free(p);
    return STP_OK;
}

int stp_memcpy_h2d(void* dst, const void* src, int64_t bytes, void*) {
    dev(dst, bytes, "stp_memcpy_h2d: destination");
    if (bytes > 0) {
        host(src, "stp_memcpy_h2d: source is not host memory");
        std::memcpy(dst, src, static_cast<size_t>(bytes));
    }
    return STP_OK;
}

int stp_memcpy_d2h(void* dst, const void* src, int64_t bytes, void*) {
    dev(src, bytes, "stp_memcpy_d2h: source");
    if (bytes > 0) {
        host(dst, "stp_memcpy_d2h: destination is not host memory");
        std::memcpy(dst, src, static_cast<size_t>(bytes));
    }
    return STP_OK;
}

int stp_memset_d(void* dst, int value, int64_t bytes, void*) {
    dev(dst, bytes, "stp_memset_d");
    if (bytes > 0) std::memset(dst, value, static_cast<size_t>(bytes));
    return STP_OK;
}

int stp_stream_sync(void*) { return STP_OK; }

int stp_csr_assemble(int64_t rows, int64_t cols, int64_t n, const int64_t* r, const int64_t* c, const double* v,
                     int64_t* rp, int64_t* ci, double* val, int64_t* nnz, void*) {
    if (rows < 0 || cols < 0 || n < 0) return fail(STP_ERR_INVALID, "stp_csr_assemble: bad shape");
    dev(r, 8 * n, "stp_csr_assemble: r");
    dev(c, 8 * n, "stp_csr_assemble: c");
    dev(v, 8 * n, "stp_csr_assemble: v");
    dev(rp, 8 * (rows + 1), "stp_csr_assemble: row_ptr");
    dev(ci, 8 * n, "stp_csr_assemble: col_out");
    dev(val, 8 * n, "stp_csr_assemble: val_out");
    host(nnz, "stp_csr_assemble: nnz");
    *nnz = orc_csr_from_triplets(rows, cols, n, r, c, v, rp, ci, val);
    return STP_OK;
}

int stp_csr64_fd7_triplets(int64_t nx, int64_t ny, int64_t nz, const double* vel, double dx, double dy, double dz,
                           double dt, int64_t* r, int64_t* c, double* v, void*) {
    if (nx < 1 || ny < 1 || nz < 1) return fail(STP_ERR_INVALID, "stp_csr64_fd7_triplets: bad shape");
    const int64_t n = nx * ny * nz;
    dev(vel, 8 * n, "stp_csr64_fd7_triplets: velocity");
    dev(r, 56 * n, "stp_csr64_fd7_triplets: r");
    dev(c, 56 * n, "stp_csr64_fd7_triplets: c");
    dev(v, 56 * n, "stp_csr64_fd7_triplets: v");
    // 7 slots per cell: centre, then x-, x+, y-, y+, z-, z+ (row -1 when absent)
    const double ix2 = 1.0 / (dx * dx), iy2 = 1.0 / (dy * dy), iz2 = 1.0 / (dz * dz);
    for (int64_t i = 0; i < n; ++i) {
        const int64_t ix = i / (ny * nz), iy = (i / nz) % ny, iz = i % nz;
        const double w = vel[i] * vel[i] * dt * dt;
        const int64_t nb[6] = {i - ny * nz, i + ny * nz, i - nz, i + nz, i - 1, i + 1};
        const bool ok[6] = {ix > 0, ix < nx - 1, iy > 0, iy < ny - 1, iz > 0, iz < nz - 1};
        const double co[6] = {ix2, ix2, iy2, iy2, iz2, iz2};
        r[7 * i] = i;
        c[7 * i] = i;
        v[7 * i] = -2.0 * w * (ix2 + iy2 + iz2) + 2.0;
        for (int k = 0; k < 6; ++k) {
            r[7 * i + 1 + k] = ok[k] ? i : -1;
            c[7 * i + 1 + k] = nb[k];
            v[7 * i + 1 + k] = w * co[k];
        }
    }
    return STP_OK;
}

int stp_csr64_to_f32(int64_t rows, int64_t nnz, const int64_t* rp, const int64_t* ci, const double* val,
                     int32_t* rp32, int32_t* ci32, float* val32, void*) {
    if (rows < 0 || nnz < 0 || rows > 0x7ffffffe || nnz > 0x7fffffff)
        return fail(STP_ERR_INVALID, "stp_csr64_to_f32: bad shape");
    dev(rp, 8 * (rows + 1), "stp_csr64_to_f32: row_ptr");
    dev(ci, 8 * nnz, "stp_csr64_to_f32: col");
    dev(val, 8 * nnz, "stp_csr64_to_f32: val");
    dev(rp32, 4 * (rows + 1), "stp_csr64_to_f32: row_ptr32");
    dev(ci32, 4 * nnz, "stp_csr64_to_f32: col32");
    dev(val32, 4 * nnz, "stp_csr64_to_f32: val32");
    for (int64_t i = 0; i <= rows; ++i) rp32[i] = static_cast<int32_t>(rp[i]);
    for (int64_t k = 0; k < nnz; ++k) {
        ci32[k] = static_cast<int32_t>(ci[k]);
        val32[k] = static_cast<float>(val[k]);
    }
    return STP_OK;
}

int stp_csr_create(int64_t rows, int64_t cols, int64_t nnz, const int32_t* rp, const int32_t* ci, const float* val,
                   stp_csr** out, void*) {
    host(out, "stp_csr_create: result pointer");
    if (rows < 0 || cols < 0 || nnz < 0 || rows > 0x7ffffffe || cols > 0x7ffffffe || nnz > 0x7fffffff)
        return fail(STP_ERR_INVALID, "stp_csr_create: bad shape");
    dev(rp, 4 * (rows + 1), "stp_csr_create: row_ptr");
    dev(ci, 4 * nnz, "stp_csr_create: col");
    dev(val, 4 * nnz, "stp_csr_create: val");
    if (rp[0] != 0 || rp[rows] != nnz) die("stp_csr_create: row_ptr does not run from 0 to nnz", rp, rows + 1);
    for (int64_t i = 0; i < rows; ++i)
        if (rp[i + 1] < rp[i]) die("stp_csr_create: row_ptr decreases", rp + i, 8);
    for (int64_t k = 0; k < nnz; ++k)
        if (ci[k] < 0 || ci[k] >= cols) die("stp_csr_create: column out of range", ci + k, 4);
    auto* A = new stp_csr();
    A->rows = rows;
    A->cols = cols;
    A->nnz = nnz;
    A->rp.assign(rp, rp + rows + 1);
    A->ci.assign(ci, ci + nnz);
    A->v32.assign(val, val + nnz);
    *out = A;
    return STP_OK;
}

int stp_csr_destroy(stp_csr* A) {
    delete A;
    return STP_OK;
}

int stp_csr_set_values_f64(stp_csr* A, const double* v, void*) {
    if (!A) return fail(STP_ERR_INVALID, "stp_csr_set_values_f64: null matrix");
    dev(v, 8 * A->nnz, "stp_csr_set_values_f64: values");
    A->v64.assign(v, v + A->nnz);
    return STP_OK;
}

int stp_csr_set_grid(stp_csr* A, int64_t nx, int64_t ny, int64_t nz) {
    if (!A) return fail(STP_ERR_INVALID, "stp_csr_set_grid: null matrix");
    if (!(nx == 0 && ny == 0 && nz == 0) && (nx < 1 || ny < 1 || nz < 1 || nx * ny * nz != A->cols))
        return fail(STP_ERR_INVALID, "stp_csr_set_grid: grid does not match the columns");
    A->grid[0] = nx;
    A->grid[1] = ny;
    A->grid[2] = nz;
    return STP_OK;
}

// FP32 matrix values, FP64 vectors and accumulation, as the library's product
int stp_csr_spmv_f64(const stp_csr* A, int32_t transpose, const double* x, double* y, void*) {
    if (!A) return fail(STP_ERR_INVALID, "stp_csr_spmv_f64: null matrix");
    const int64_t nin = transpose ? A->rows : A->cols, nout = transpose ? A->cols : A->rows;
    dev(x, 8 * nin, "stp_csr_spmv_f64: x");
    dev(y, 8 * nout, "stp_csr_spmv_f64: y");
    if (transpose)
        orc_csr_spmv_t(A->rows, A->cols, A->rp.data(), A->ci.data(), A->v32.data(), x, y);
    else
        orc_csr_spmv(A->rows, A->rp.data(), A->ci.data(), A->v32.data(), x, y);
    return STP_OK;
}

int stp_dot_f64(int64_t n, const double* a, const double* b, double* out, void*) {
    if (n < 0) return fail(STP_ERR_INVALID, "stp_dot_f64: bad size");
    dev(a, 8 * n, "stp_dot_f64: a");
    dev(b, 8 * n, "stp_dot_f64: b");
    host(out, "stp_dot_f64: result is not host memory");
    *out = orc_dot(a, b, n);
    return STP_OK;
}

int stp_cg(const stp_csr* A, const double* b, int32_t max_iter, double tol, int32_t jacobi, double* x,
           double* history, stp_solve_info* info, void*) {
    if (!A || max_iter < 0) return fail(STP_ERR_INVALID, "stp_cg: bad argument");
    if (A->rows != A->cols) return fail(STP_ERR_INVALID, "stp_cg: matrix is not square");
    const int64_t n = A->rows;
    dev(b, 8 * n, "stp_cg: b");
    dev(x, 8 * n, "stp_cg: x");
    if (history) dev(history, 8 * max_iter, "stp_cg: history");
    const std::vector<double>& v = values(A);
    double rel = 0.0;
    const int it = jacobi ? orc_jacobi_pcg(n, A->rp.data(), A->ci.data(), v.data(), b, max_iter, tol, x, &rel)
                          : orc_cg(n, A->rp.data(), A->ci.data(), v.data(), b, max_iter, tol, x, &rel);
    std::vector<double> ax(static_cast<size_t>(n));
    orc_csr_spmv(n, A->rp.data(), A->ci.data(), v.data(), x, ax.data());
    double rr = 0.0, bb = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        rr += (b[i] - ax[i]) * (b[i] - ax[i]);
        bb += b[i] * b[i];
    }
    if (info) {
        host(info, "stp_cg: info is not host memory");
        info->iterations = it;
        info->rel_residual = rel;
        info->true_rel_residual = bb > 0 ? std::sqrt(rr / bb) : 0.0;
        info->converged = info->true_rel_residual < tol;
        info->replacements = 0;
    }
    return STP_OK;
}

int stp_cgls(const stp_csr* A, const double* b, double lam, double tol, int32_t max_iter, int32_t precond,
             void* precond_handle, double* x, double* history, stp_solve_info* info, void*) {
    if (!A || max_iter < 0 || lam < 0 || precond < 0 || precond > 2)
        return fail(STP_ERR_INVALID, "stp_cgls: bad argument");
    if (precond == 2 || precond_handle) return fail(STP_ERR_STATE, "stp_cgls: no learned preconditioner here");
    dev(b, 8 * A->rows, "stp_cgls: b");
    dev(x, 8 * A->cols, "stp_cgls: x");
    if (history) dev(history, 8 * max_iter, "stp_cgls: history");
    const std::vector<double>& v = values(A);
    double rel = 0.0;
    const int it = orc_cgls(A->rows, A->cols, A->rp.data(), A->ci.data(), v.data(), b, lam, tol, max_iter, precond, x,
                            &rel, history);
    if (info) {
        host(info, "stp_cgls: info is not host memory");
        info->iterations = it;
        info->rel_residual = rel;
        info->true_rel_residual = rel;
        info->converged = rel < tol;
        info->replacements = 0;
    }
    return STP_OK;
}

int stp_smooth_gradient_f64(int64_t nx, int64_t ny, int64_t nz, double sigma, const double* in, double* out,
                            double* work, void*) {
    if (nx < 1 || ny < 1 || nz < 1 || !(sigma > 0)) return fail(STP_ERR_INVALID, "stp_smooth_gradient_f64: bad argument");
    const int64_t n = nx * ny * nz;
    dev(in, 8 * n, "stp_smooth_gradient_f64: in");
    dev(out, 8 * n, "stp_smooth_gradient_f64: out");
    dev(work, 8 * n, "stp_smooth_gradient_f64: work");
    orc_smooth_gradient(nx, ny, nz, sigma, in, out);
    return STP_OK;
}

}  // extern "C"
