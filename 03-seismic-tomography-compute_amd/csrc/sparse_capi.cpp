// C ABI of the tomography solve (declarations: include/seismic_hip.h).
//
// stp_csr holds A (FP32 values, int32 columns/row offsets) and an explicit
// A^T built on the device, plus a lazily sized solver workspace.  Solvers
// keep their scalars on the device and only synchronise with the host every
// CHECK iterations to read the convergence flag.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "capi_common.h"
#include "csr_handle.h"
#include "learned_precond.h"
#include "seismic_hip.h"
#include "solver_layout.h"
#include "sparse_kernels.h"

using stp::fail;
using stp::hip_fail;
using stp::S;

namespace {
constexpr int CHECK = 8;   // iterations between host reads of the done flag
// CGLS: the smallest tolerance solved with r gathered in FP32 for A^T r, and
// the iterations without a 10 % drop of the residual after which such a pass
// counts as stalled
constexpr double R32_MIN_TOL = 1e-6;
constexpr int STALL = 512;

template <typename T>
hipError_t grow(T** p, size_t* cap, size_t n) {
    if (*cap >= n && *p) return hipSuccess;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * (n ? n : 1));
    if (e == hipSuccess) *cap = n;
    return e;
}
}  // namespace


static void csr_free(stp_csr* h) {
    void* ps[] = {h->rp, h->ci, h->trp, h->tci, h->val, h->tval, h->val64, h->r, h->part0, h->part1, h->part2, h->part3, h->part4, h->part5,
                  h->p, h->q, h->s, h->z, h->sp, h->dinv, h->r32, h->ry, h->rg, h->pnorm, h->dp, h->dap, h->dz, h->dr, h->ddinv, h->sc};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    stp::free_solver_layout(h);
}

static int finish_csr(stp_csr* h, hipStream_t st) {
    STP_HIP(hipMalloc(&h->trp, sizeof(int) * ((size_t)h->cols + 1)));
    STP_HIP(hipMalloc(&h->tci, sizeof(int) * (h->nnz ? h->nnz : 1)));
    STP_HIP(hipMalloc(&h->tval, sizeof(float) * (h->nnz ? h->nnz : 1)));
    STP_HIP(stp::csr_transpose(h->rows, h->cols, h->nnz, h->rp, h->ci, h->val, h->trp, h->tci, h->tval, st));
    h->W = stp::spmv_lanes(h->rows ? (double)h->nnz / h->rows : 0.0, 0);
    h->WT = stp::spmv_lanes(h->cols ? (double)h->nnz / h->cols : 0.0, 1);
    return STP_OK;
}

extern "C" {

int stp_csr_create(int64_t rows, int64_t cols, int64_t nnz, const int32_t* rp, const int32_t* ci, const float* val,
                   stp_csr** out, void* stream) {
    if (!out || rows < 0 || cols < 0 || nnz < 0 || rows > 0x7ffffffe || cols > 0x7ffffffe || nnz > 0x7fffffff ||
        !rp || (nnz > 0 && (!ci || !val)))
        return fail(STP_ERR_INVALID, "stp_csr_create: bad argument");
    *out = nullptr;
    stp_csr* h = new (std::nothrow) stp_csr();
    if (!h) return fail(STP_ERR_ALLOC, "stp_csr_create: out of host memory");
    h->rows = (int)rows; h->cols = (int)cols; h->nnz = nnz;
    hipStream_t st = S(stream);
    hipError_t e;
    if ((e = hipMalloc(&h->rp, sizeof(int) * (rows + 1))) != hipSuccess ||
        (e = hipMalloc(&h->ci, sizeof(int) * (nnz ? nnz : 1))) != hipSuccess ||
        (e = hipMalloc(&h->val, sizeof(float) * (nnz ? nnz : 1))) != hipSuccess ||
        (e = hipMemcpyAsync(h->rp, rp, sizeof(int) * (rows + 1), hipMemcpyDeviceToDevice, st)) != hipSuccess ||
        (nnz && (e = hipMemcpyAsync(h->ci, ci, sizeof(int) * nnz, hipMemcpyDeviceToDevice, st)) != hipSuccess) ||
        (nnz && (e = hipMemcpyAsync(h->val, val, sizeof(float) * nnz, hipMemcpyDeviceToDevice, st)) != hipSuccess)) {
        csr_free(h);
        delete h;
        return hip_fail(e, "stp_csr_create");
    }
    const int rc = finish_csr(h, st);
    if (rc != STP_OK) { csr_free(h); delete h; return rc; }
    *out = h;
    return STP_OK;
}

int stp_csr_from_rays(int64_t nrays, const int64_t* n, double h_cell, const double* pts, stp_csr** out,
                      void* stream) {
    if (!out || !n || nrays < 0 || nrays > 0x7ffffffe || (nrays > 0 && !pts) || !(h_cell > 0) || n[0] < 1 ||
        n[1] < 1 || n[2] < 1 || n[0] * n[1] * n[2] > 0x7ffffffe)
        return fail(STP_ERR_INVALID, "stp_csr_from_rays: bad argument");
    *out = nullptr;
    hipStream_t st = S(stream);
    stp_csr* h = new (std::nothrow) stp_csr();
    if (!h) return fail(STP_ERR_ALLOC, "stp_csr_from_rays: out of host memory");
    h->rows = (int)nrays;
    h->cols = (int)(n[0] * n[1] * n[2]);
    for (int k = 0; k < 3; ++k) h->gn[k] = n[k];   // columns are grid cells: CGLS uses the solver layout
    int* counts = nullptr;
    hipError_t e = hipSuccess;
    int last_count = 0, last_off = 0;
    do {
        if ((e = hipMalloc(&counts, sizeof(int) * (nrays + 1))) != hipSuccess) break;
        if ((e = hipMemsetAsync(counts, 0, sizeof(int) * (nrays + 1), st)) != hipSuccess) break;
        if ((e = stp::rays_count(nrays, pts, n, h_cell, counts, st)) != hipSuccess) break;
        if ((e = hipMalloc(&h->rp, sizeof(int) * (nrays + 1))) != hipSuccess) break;
        if ((e = stp::exclusive_scan_i32(counts, h->rp, (int)nrays + 1, st)) != hipSuccess) break;
        if ((e = hipMemcpy(&last_off, h->rp + nrays, sizeof(int), hipMemcpyDeviceToHost)) != hipSuccess) break;
        (void)last_count;
        h->nnz = last_off;
        if ((e = hipMalloc(&h->ci, sizeof(int) * (h->nnz ? h->nnz : 1))) != hipSuccess) break;
        if ((e = hipMalloc(&h->val, sizeof(float) * (h->nnz ? h->nnz : 1))) != hipSuccess) break;
        if ((e = stp::rays_fill(nrays, pts, n, h_cell, h->rp, h->ci, h->val, st)) != hipSuccess) break;
        e = hipStreamSynchronize(st);
    } while (0);
    if (counts) (void)hipFree(counts);
    if (e != hipSuccess) { csr_free(h); delete h; return hip_fail(e, "stp_csr_from_rays"); }
    const int rc = finish_csr(h, st);
    if (rc != STP_OK) { csr_free(h); delete h; return rc; }
    *out = h;
    return STP_OK;
}

int stp_csr_destroy(stp_csr* h) {
    if (!h) return STP_OK;
    csr_free(h);
    delete h;
    return STP_OK;
}

int stp_csr_shape(const stp_csr* h, int64_t* rows, int64_t* cols, int64_t* nnz) {
    if (!h || !rows || !cols || !nnz) return fail(STP_ERR_INVALID, "stp_csr_shape: null argument");
    *rows = h->rows; *cols = h->cols; *nnz = h->nnz;
    return STP_OK;
}

int stp_csr_export(const stp_csr* h, int32_t transpose, int32_t* rp, int32_t* ci, float* val, void* stream) {
    if (!h || !rp || (h->nnz && (!ci || !val))) return fail(STP_ERR_INVALID, "stp_csr_export: null argument");
    hipStream_t st = S(stream);
    const int n = transpose ? h->cols : h->rows;
    STP_HIP(hipMemcpyAsync(rp, transpose ? h->trp : h->rp, sizeof(int) * ((size_t)n + 1), hipMemcpyDeviceToDevice, st));
    if (h->nnz) {
        STP_HIP(hipMemcpyAsync(ci, transpose ? h->tci : h->ci, sizeof(int) * h->nnz, hipMemcpyDeviceToDevice, st));
        STP_HIP(hipMemcpyAsync(val, transpose ? h->tval : h->val, sizeof(float) * h->nnz, hipMemcpyDeviceToDevice, st));
    }
    return STP_OK;
}

}  // extern "C"

// y = A x / A^T x through the solver layout (when it is built): x is permuted
// into layout order, multiplied by the renumbered copy and the result
// permuted back.  Temporaries come from the stream-ordered pool, so products
// on one handle may run on several streams at once.
template <typename T>
static int layout_spmv(const stp_csr* h, int transpose, const T* x, T* y, hipStream_t st) {
    const int rows = h->rows, cols = h->cols;
    const int nin = transpose ? rows : cols, nout = transpose ? cols : rows;
    T *xl = nullptr, *yl = nullptr;
    STP_HIP(hipMallocAsync(reinterpret_cast<void**>(&xl), sizeof(T) * (size_t)(nin > 0 ? nin : 1), st));
    hipError_t e = hipMallocAsync(reinterpret_cast<void**>(&yl), sizeof(T) * (size_t)(nout > 0 ? nout : 1), st);
    if (e == hipSuccess) {
        if (!transpose) {   // x'[cperm[j]] = x[j]; y[rsrc[k]] = (A' x')[k]
            e = stp::permute<T>(cols, x, h->cperm, xl, true, st);
            if (e == hipSuccess) {
                if constexpr (sizeof(T) == 4)
                    e = stp::spmv_f32(h->lW, rows, h->lrp, h->lci, h->lval, xl, yl, nullptr, nullptr, st);
                else
                    e = stp::spmv_f64(h->lW, rows, h->lrp, h->lci, h->lval, xl, yl, 1.0, nullptr, 0.0, nullptr,
                                      nullptr, st);
            }
            if (e == hipSuccess) e = stp::permute<T>(rows, yl, h->rsrc, y, true, st);
        } else {            // x'[k] = x[rsrc[k]]; y[j] = (A'^T x')[cperm[j]]
            e = stp::permute<T>(rows, x, h->rsrc, xl, false, st);
            if (e == hipSuccess) {
                if constexpr (sizeof(T) == 4)
                    e = stp::spmv_f32(h->lWT, cols, h->ltrp, h->ltci, h->ltval, xl, yl, nullptr, nullptr, st);
                else
                    e = stp::spmv_f64(h->lWT, cols, h->ltrp, h->ltci, h->ltval, xl, yl, 1.0, nullptr, 0.0, nullptr,
                                      nullptr, st);
            }
            if (e == hipSuccess) e = stp::permute<T>(cols, yl, h->cperm, y, false, st);
        }
    }
    const hipError_t e1 = hipFreeAsync(xl, st);
    const hipError_t e2 = yl ? hipFreeAsync(yl, st) : hipSuccess;
    if (e == hipSuccess) e = e1 != hipSuccess ? e1 : e2;
    if (e != hipSuccess) return stp::hip_fail(e, "sparse product in the solver layout");
    return STP_OK;
}

extern "C" {

int stp_csr_spmv(const stp_csr* h, int32_t transpose, const float* x, float* y, void* stream) {
    if (!h || !x || !y) return fail(STP_ERR_INVALID, "stp_csr_spmv: null argument");
    if (h->lay_ready) return layout_spmv<float>(h, transpose, x, y, S(stream));
    if (transpose)
        STP_HIP(stp::spmv_f32(h->WT, h->cols, h->trp, h->tci, h->tval, x, y, nullptr, nullptr, S(stream)));
    else
        STP_HIP(stp::spmv_f32(h->W, h->rows, h->rp, h->ci, h->val, x, y, nullptr, nullptr, S(stream)));
    return STP_OK;
}

int stp_csr_spmv_f64(const stp_csr* h, int32_t transpose, const double* x, double* y, void* stream) {
    if (!h || !x || !y) return fail(STP_ERR_INVALID, "stp_csr_spmv_f64: null argument");
    if (h->lay_ready) return layout_spmv<double>(h, transpose, x, y, S(stream));
    if (transpose)
        STP_HIP(stp::spmv_f64(h->WT, h->cols, h->trp, h->tci, h->tval, x, y, 1.0, nullptr, 0.0, nullptr, nullptr,
                              S(stream)));
    else
        STP_HIP(stp::spmv_f64(h->W, h->rows, h->rp, h->ci, h->val, x, y, 1.0, nullptr, 0.0, nullptr, nullptr,
                              S(stream)));
    return STP_OK;
}

// Measurement: the gathers and index stream of one product through the
// matrix the solver uses (the solver layout when built, else the given
// order), without the value stream.
int stp_csr_gather_probe(const stp_csr* h, int32_t transpose, int32_t fp64, const void* x, void* y, void* stream) {
    if (!h || !x || !y) return fail(STP_ERR_INVALID, "stp_csr_gather_probe: null argument");
    const bool lay = h->lay_ready;
    const int n = transpose ? h->cols : h->rows;
    const int* rp = transpose ? (lay ? h->ltrp : h->trp) : (lay ? h->lrp : h->rp);
    const int* ci = transpose ? (lay ? h->ltci : h->tci) : (lay ? h->lci : h->ci);
    const int W = transpose ? (lay ? h->lWT : h->WT) : (lay ? h->lW : h->W);
    hipStream_t st = S(stream);
    if (fp64)
        STP_HIP(stp::gather_probe<double>(W, n, rp, ci, static_cast<const double*>(x), static_cast<double*>(y), st));
    else
        STP_HIP(stp::gather_probe<float>(W, n, rp, ci, static_cast<const float*>(x), static_cast<float*>(y), st));
    return STP_OK;
}

// ------------------------------------------------------------------ square CG / Jacobi-PCG
int stp_csr_set_values_f64(stp_csr* h, const double* val64, void* stream) {
    if (!h) return fail(STP_ERR_INVALID, "stp_csr_set_values_f64: null handle");
    if (!val64) {
        if (h->val64) (void)hipFree(h->val64);
        h->val64 = nullptr;
        return STP_OK;
    }
    if (!h->val64) STP_HIP(hipMalloc(&h->val64, sizeof(double) * (h->nnz ? h->nnz : 1)));
    if (h->nnz)
        STP_HIP(hipMemcpyAsync(h->val64, val64, sizeof(double) * h->nnz, hipMemcpyDeviceToDevice, S(stream)));
    return STP_OK;
}

// An event owned for the duration of one solve.
struct OwnedEvent {
    hipEvent_t ev = nullptr;
    OwnedEvent() {
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) ev = nullptr;
    }
    ~OwnedEvent() {
        if (ev) (void)hipEventDestroy(ev);
    }
    OwnedEvent(const OwnedEvent&) = delete;
    OwnedEvent& operator=(const OwnedEvent&) = delete;
};

int stp_cg(const stp_csr* hc, const double* b, int32_t max_iter, double tol, int32_t jacobi, double* x,
           double* history, stp_solve_info* info, void* stream) {
    stp_csr* h = const_cast<stp_csr*>(hc);
    if (!h || !b || !x || max_iter < 0) return fail(STP_ERR_INVALID, "stp_cg: bad argument");
    if (h->rows != h->cols) return fail(STP_ERR_INVALID, "stp_cg: matrix is not square");
    hipStream_t st = S(stream);
    const int n = h->rows;
    const int nbd = stp::cg_blocks(n);
    const double* v64 = h->val64;
    STP_HIP(grow(&h->dp, &h->c_dp, n));
    STP_HIP(grow(&h->dap, &h->c_dap, n));
    STP_HIP(grow(&h->dz, &h->c_dz, n));
    STP_HIP(grow(&h->dr, &h->c_dr, n));
    STP_HIP(grow(&h->ddinv, &h->c_ddinv, n));
    STP_HIP(grow(&h->part0, &h->c_p0, nbd));
    STP_HIP(grow(&h->part1, &h->c_p1, nbd));
    STP_HIP(grow(&h->sc, &h->c_sc, 1));
    double* dinv = nullptr;
    if (jacobi) {
        if (v64) STP_HIP(stp::launch_diag_inv_f64v(n, h->rp, h->ci, v64, h->ddinv, st));
        else STP_HIP(stp::launch_diag_inv(n, h->rp, h->ci, h->val, h->ddinv, st));
        dinv = h->ddinv;
    }
    // A p with the most precise values the handle holds
    auto matvec = [&](const double* in, double* out, const int* done) -> hipError_t {
        if (v64) return stp::spmv_f64v(h->W, n, h->rp, h->ci, v64, in, out, done, st);
        return stp::spmv_f64(h->W, n, h->rp, h->ci, h->val, in, out, 1.0, nullptr, 0.0, nullptr, done, st);
    };
    double* z = jacobi ? h->dz : h->dr;
    STP_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, st));
    STP_HIP(hipMemcpyAsync(h->dr, b, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    if (jacobi) {
        // z = D^-1 r and the r.r / r.z partials through the update kernel with alpha = 0
        STP_HIP(hipMemsetAsync(h->sc, 0, sizeof(stp::SolverScalars), st));
        STP_HIP(hipMemsetAsync(h->dap, 0, sizeof(double) * n, st));
        STP_HIP(stp::launch_cg_update(h->sc, n, x, h->dr, h->dap, h->dap, dinv, h->dz, h->part0, h->part1, st));
    } else {
        STP_HIP(stp::launch_dot_partial(nullptr, n, h->dr, h->dr, h->part0, st));
        STP_HIP(stp::launch_dot_partial(nullptr, n, h->dr, h->dr, h->part1, st));
    }
    STP_HIP(hipMemcpyAsync(h->dp, z, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    STP_HIP(stp::launch_cg_init(h->sc, h->part0, h->part1, nbd, st));
    int* done = &h->sc->done;
    int it = 0;
    stp::SolverScalars hs;
    std::memset(&hs, 0, sizeof hs);
    OwnedEvent ev;
    bool ap_queued = false;   // this iteration's A p was queued before the last host read
    for (; it < max_iter; ++it) {
        if (!ap_queued) STP_HIP(matvec(h->dp, h->dap, done));
        ap_queued = false;
        STP_HIP(stp::launch_dot_partial(done, n, h->dp, h->dap, h->part0, st));
        STP_HIP(stp::launch_cg_alpha(h->sc, h->part0, nbd, st));
        STP_HIP(stp::launch_cg_update(h->sc, n, x, h->dr, h->dp, h->dap, dinv, h->dz, h->part0, h->part1, st));
        STP_HIP(stp::launch_cg_beta(h->sc, h->part0, h->part1, nbd, tol, it, history, st));
        STP_HIP(stp::launch_cg_direction(h->sc, n, z, h->dp, st));
        if ((it + 1) % CHECK == 0 || it + 1 == max_iter) {
            STP_HIP(hipMemcpyAsync(&hs, h->sc, sizeof hs, hipMemcpyDeviceToHost, st));
            if (ev.ev && it + 1 < max_iter) {
                // the next A p is queued before the host waits (after
                // convergence it returns at once on the done flag)
                STP_HIP(hipEventRecord(ev.ev, st));
                STP_HIP(matvec(h->dp, h->dap, done));
                ap_queued = true;
                STP_HIP(hipEventSynchronize(ev.ev));
            } else {
                STP_HIP(hipStreamSynchronize(st));
            }
            if (hs.done) break;
        }
    }
    STP_HIP(hipMemcpyAsync(&hs, h->sc, sizeof hs, hipMemcpyDeviceToHost, st));
    // true residual b - A x (into dap: the iteration no longer needs it)
    if (v64) STP_HIP(stp::residual_f64v(h->W, n, h->rp, h->ci, v64, x, b, h->dap, st));
    else STP_HIP(stp::residual(h->W, n, h->rp, h->ci, h->val, x, b, h->dap, st));
    STP_HIP(stp::launch_dot_partial(nullptr, n, h->dap, h->dap, h->part0, st));
    std::vector<double> pr(nbd);
    STP_HIP(hipMemcpyAsync(pr.data(), h->part0, sizeof(double) * nbd, hipMemcpyDeviceToHost, st));
    STP_HIP(hipStreamSynchronize(st));
    double rr = 0.0;
    for (double v : pr) rr += v;
    if (info) {
        info->iterations = hs.iters;
        info->converged = hs.done == 1;
        info->rel_residual = hs.rel;
        info->true_rel_residual = hs.s0 > 1e-30 ? std::sqrt(rr) / hs.s0 : 0.0;
        info->replacements = 0;
    }
    return STP_OK;
}

// ------------------------------------------------------------------ CGLS
// CGLS on (A^T A + lam I) x = A^T b, x0 = 0.  precond 0: none; 1: Jacobi
// (1 / column sums of squares + lam); 2: learned preconditioner (prepared by
// stp_lpc_prepare), with the flexible beta = z.(s - s_prev) / gamma_prev,
// which equals the standard one for a fixed linear M and tolerates its
// BF16-derived coefficients.  Vectors p, q, s, z are FP32 (the SpMV
// A p accumulates in FP32); x and r are FP64 and every dot product is FP64.
// When the recursive residual reaches tol the true residual is recomputed in
// FP64 (r = b - A x, s = A^T r - lam x); if it is still above tol the
// iteration restarts from it (iterative refinement of the FP32 inner solve).
// A matrix whose columns are grid cells (stp_csr_from_rays, stp_csr_set_grid)
// is solved in its solver layout: b is gathered into layout row order, the
// iteration runs on the renumbered copy and x is gathered back to cell order.
}  // extern "C"

namespace {

// Device workspace of one CGLS solve: r (rows, FP64), q (rows, FP32), p, s,
// z, s_prev and the Jacobi diagonal (cols, FP32), for a pass that gathers
// FP32 the change y of r (rows, FP64), the FP32 copy of the gathered vector
// (rows) and g = A^T r at the start of the pass (cols, FP64), the per-block
// partial sums and the device-resident scalars.
struct CglsWork {
    double *r = nullptr, *y = nullptr, *g = nullptr;
    float *p = nullptr, *q = nullptr, *s = nullptr, *z = nullptr, *sp = nullptr, *dinv = nullptr, *v32 = nullptr;
    double *part0 = nullptr, *part1 = nullptr, *part2 = nullptr, *part3 = nullptr, *part4 = nullptr,
           *part5 = nullptr, *pnorm = nullptr;   // pnorm: |r|^2 and |y|^2 partials, 2 x 256
    stp::SolverScalars* sc = nullptr;
    // hipEvent_t for the host reads of the scalars (null: the reads synchronize
    // the stream; the entry point that allocates nothing passes none)
    hipEvent_t ev = nullptr;
};

int cgls_check(const stp_csr* h, const double* b, double lam, int32_t max_iter, int32_t precond, void* ph,
               const double* x, const char* who) {
    if (!h || !b || !x || max_iter < 0 || lam < 0 || precond < 0 || precond > 2)
        return fail(STP_ERR_INVALID, std::string(who) + ": bad argument");
    if (precond == 2 && !ph) return fail(STP_ERR_INVALID, std::string(who) + ": learned preconditioner handle missing");
    if (precond == 2 && stp::lpc_cols(ph) != h->cols)
        return fail(STP_ERR_STATE, std::string(who) + ": learned preconditioner is not prepared for this matrix");
    return STP_OK;
}

// The solve itself: kernel launches and host reads of the scalars only (the
// solver layout, when A has a grid, must be built).
int cgls_solve(stp_csr* h, const double* b, double lam, double tol, int32_t max_iter, int32_t precond,
               void* precond_handle, double* x, double* history, stp_solve_info* info, hipStream_t st,
               const CglsWork& w) {
    const int rows = h->rows, cols = h->cols;
    const bool lay = h->gn[0] > 0;
    const int* RP = lay ? h->lrp : h->rp;
    const int* CI = lay ? h->lci : h->ci;
    const float* VAL = lay ? h->lval : h->val;
    const int* TRP = lay ? h->ltrp : h->trp;
    const int* TCI = lay ? h->ltci : h->tci;
    const float* TVAL = lay ? h->ltval : h->tval;
    const int* perm = lay ? h->cperm : nullptr;
    const int W = lay ? h->lW : h->W, WT = lay ? h->lWT : h->WT;
    double* const x_out = x;
    if (lay) {
        STP_HIP(stp::permute<double>(rows, b, h->rsrc, h->lb, false, st));
        b = h->lb;
        x = h->lx;
    }
    const int nbq = std::max(stp::spmv_blocks(rows, W), 1);
    const int nbs = std::max(stp::spmv_blocks(cols, WT), 1);
    const int nbc = std::max((cols + 255) / 256, 1);   // partials of the learned apply
    const int npp = stp::cgls_blocks(cols);            // |p|^2 partials of the direction kernel
    const bool flex = precond == 2;
    float* dinv = nullptr;
    if (precond == 1) {
        STP_HIP(stp::launch_normal_diag(cols, TRP, TVAL, lam, w.dinv, st));
        dinv = w.dinv;
    }
    int* done = &w.sc->done;
    const size_t fw = sizeof(float) / 4, dw = sizeof(double) / 4;
    // For tol >= R32_MIN_TOL the iteration's A^T r gathers an FP32 vector (in
    // the solver layout from the staged A^T, whose tiles read their distinct
    // rays once, into LDS; FP64 sums).  Its rounding error is relative to the
    // vector gathered, so a pass starts with g - A^T y, g = A^T r_ref in FP64
    // at the start of the pass and y = r_ref - r (the sum of the alpha q since
    // then), and switches to A^T r itself once |y| exceeds |r| (checked every
    // CHECK iterations): consistent data drive r to 0, data with a large part
    // outside A's range keep |r| large while |y| stays bounded by the part in
    // range.  r is updated in FP64 throughout.  A pass that stalls (no 10 %
    // gain in STALL iterations) or ends above tol in the FP64 check restarts
    // from the recomputed residual with r in FP64.  Odd iterations sweep the
    // rows of both products last to first, so each product starts on the part
    // of its matrix that the previous sweep left in the device cache.
    bool use32 = tol >= R32_MIN_TOL;
    const int npa = stp::cgls_blocks(cols > rows ? cols : rows);   // blocks of the alpha/update kernel
    double hnorm[2 * 256];   // npa <= 256

    // s = A^T r - lam x, z = M s, gamma = s.z, p = z, s_prev = s (and g, y = 0)
    auto start = [&](int set_s0, int it0) -> int {
        STP_HIP(stp::normal_residual(WT, cols, TRP, TCI, TVAL, w.r, x, lam, dinv, w.s, w.z, w.part2, w.part3, nullptr,
                                     st, 0, use32 ? w.g : nullptr));
        if (use32) STP_HIP(stp::launch_zero_words(w.y, dw * rows, st));
        if (flex) {
            STP_HIP(stp::launch_zero_words(w.sp, fw * cols, st));
            int rc = stp::lpc_apply(precond_handle, w.s, w.z, nullptr, w.sp, w.part4, w.part5, st, perm);
            if (rc != STP_OK) return rc;
            STP_HIP(stp::launch_cgls_init(w.sc, w.part4, nbc, w.part3, nbs, set_s0, it0, st));
        } else {
            STP_HIP(stp::launch_cgls_init(w.sc, w.part2, nbs, w.part3, nbs, set_s0, it0, st));
        }
        STP_HIP(stp::launch_zero_words(w.p, fw * cols, st));
        STP_HIP(stp::launch_cgls_beta_direction(w.sc, nullptr, 0, nullptr, 0, nullptr, 0, tol, nullptr, it0, 1, cols,
                                                w.z, w.p, w.part1, w.s, w.sp, st));
        return STP_OK;
    };

    STP_HIP(stp::launch_zero_words(x, dw * cols, st));
    STP_HIP(stp::launch_zero_words(w.sc, sizeof(stp::SolverScalars) / 4, st));
    STP_HIP(hipMemcpyAsync(w.r, b, sizeof(double) * rows, hipMemcpyDeviceToDevice, st));
    int rc = start(1, 0);
    if (rc != STP_OK) return rc;
    stp::SolverScalars hs;
    std::memset(&hs, 0, sizeof hs);
    int it = 0, replacements = 0;
    double true_rel = 0.0;
    double ps[stp::SPMV_MAX_BLOCKS];
    const bool staged = lay && h->st_rt > 0;
    for (;;) {
        const int nbs_it = use32 && staged ? stp::staged_blocks(h->st_ntiles) : nbs;   // partials of A^T r
        double best = HUGE_VAL;
        int improved_at = it;
        bool stalled = false;
        bool ymode = use32;
        bool ap_queued = false;   // this iteration's A p was queued before the last host read
        for (; it < max_iter; ++it) {
            const int rev = it & 1;
            const bool check = (it + 1) % CHECK == 0 || it + 1 == max_iter;
            const bool norms = ymode && check;
            if (!ap_queued) STP_HIP(stp::spmv_f32(W, rows, RP, CI, VAL, w.p, w.q, w.part0, done, st, rev));
            ap_queued = false;
            STP_HIP(stp::launch_cgls_alpha_update(w.sc, w.part0, nbq, w.part1, npp, lam, it, cols, x, w.p, rows, w.r,
                                                  w.q, ymode ? w.y : nullptr, use32 ? w.v32 : nullptr,
                                                  norms ? w.pnorm : nullptr, norms ? w.pnorm + npa : nullptr, st));
            if (use32 && staged)
                STP_HIP(stp::normal_residual_staged(WT, cols, h->st_rt, h->st_ntiles, h->st_maxu, TRP, h->st_l16,
                                                    TVAL, h->st_tu, h->st_uniq, w.v32, ymode ? w.g : nullptr, x,
                                                    lam, dinv, w.s, w.z, w.part2, w.part3, done, rev, st));
            else if (use32)
                STP_HIP(stp::normal_residual_r32(WT, cols, TRP, TCI, TVAL, w.v32, ymode ? w.g : nullptr, x, lam,
                                                 dinv, w.s, w.z, w.part2, w.part3, done, rev, st));
            else
                STP_HIP(stp::normal_residual(WT, cols, TRP, TCI, TVAL, w.r, x, lam, dinv, w.s, w.z, w.part2, w.part3,
                                             done, st, rev));
            if (flex) {
                rc = stp::lpc_apply(precond_handle, w.s, w.z, done, w.sp, w.part4, w.part5, st, perm);
                if (rc != STP_OK) return rc;
                STP_HIP(stp::launch_cgls_beta_direction(w.sc, w.part4, nbc, w.part3, nbs_it, w.part5, nbc, tol,
                                                        history, it, 0, cols, w.z, w.p, w.part1, w.s, w.sp, st));
            } else {
                STP_HIP(stp::launch_cgls_beta_direction(w.sc, w.part2, nbs_it, w.part3, nbs_it, nullptr, 0, tol,
                                                        history, it, 0, cols, w.z, w.p, w.part1, w.s, nullptr, st));
            }
            if (check) {
                STP_HIP(hipMemcpyAsync(&hs, w.sc, sizeof hs, hipMemcpyDeviceToHost, st));
                if (norms)
                    STP_HIP(hipMemcpyAsync(hnorm, w.pnorm, sizeof(double) * 2 * npa, hipMemcpyDeviceToHost,
                                           st));
                if (w.ev && it + 1 < max_iter) {
                    // The next iteration's A p depends on nothing decided below
                    // (after convergence it returns at once, after a stall its
                    // q is not used), so it is queued before the host waits for
                    // the copies and the device is not idle meanwhile.
                    STP_HIP(hipEventRecord(w.ev, st));
                    STP_HIP(stp::spmv_f32(W, rows, RP, CI, VAL, w.p, w.q, w.part0, done, st, (it + 1) & 1));
                    ap_queued = true;
                    STP_HIP(hipEventSynchronize(w.ev));
                } else {
                    STP_HIP(hipStreamSynchronize(st));
                }
                if (hs.done) break;
                if (norms) {
                    double rr = 0.0, yy = 0.0;
                    for (int k = 0; k < npa; ++k) {
                        rr += hnorm[k];
                        yy += hnorm[npa + k];
                    }
                    if (yy > rr) ymode = false;   // from the next iteration on: A^T r
                }
                if (hs.rel < 0.9 * best) {
                    best = hs.rel;
                    improved_at = it;
                } else if (use32 && it - improved_at >= STALL) {
                    stalled = true;
                    ++it;
                    break;
                }
            }
        }
        STP_HIP(hipMemcpyAsync(&hs, w.sc, sizeof hs, hipMemcpyDeviceToHost, st));
        STP_HIP(hipStreamSynchronize(st));
        if (!stalled) it = hs.iters;
        // true residual in FP64
        STP_HIP(stp::residual(W, rows, RP, CI, VAL, x, b, w.r, st));
        STP_HIP(stp::normal_residual(WT, cols, TRP, TCI, TVAL, w.r, x, lam, dinv, w.s, w.z, w.part2, w.part3, nullptr,
                                     st));
        STP_HIP(hipMemcpyAsync(ps, w.part3, sizeof(double) * nbs, hipMemcpyDeviceToHost, st));
        STP_HIP(hipStreamSynchronize(st));
        double ss = 0.0;
        for (int k = 0; k < nbs; ++k) ss += ps[k];
        true_rel = hs.s0 > 0 ? std::sqrt(ss) / hs.s0 : 0.0;
        if (hs.done == 1 && true_rel < tol) break;
        if (it >= max_iter || replacements >= 8 || (hs.done != 1 && !stalled)) break;
        if (use32) use32 = false;   // the rest in FP64 gathers of r
        ++replacements;
        rc = start(0, it);   // cgls_init clears the done flag before the direction kernel runs
        if (rc != STP_OK) return rc;
    }
    if (lay) STP_HIP(stp::permute<double>(cols, x, h->cperm, x_out, false, st));
    if (info) {
        info->iterations = it;
        info->converged = (hs.done == 1 || hs.s0 == 0) && true_rel < tol;
        info->rel_residual = hs.rel;
        info->true_rel_residual = true_rel;
        info->replacements = replacements;
    }
    return STP_OK;
}

// ---- workspace of the no-allocation entry point: one static device array
// per GPU (hipGetSymbolAddress gives each device's copy), carved into the
// CglsWork buffers at 256-byte offsets
constexpr int64_t NA_MAX_ROWS = int64_t(1) << 20;
constexpr int64_t NA_MAX_COLS = int64_t(1) << 20;
constexpr size_t na_align(size_t b) { return (b + 255) & ~size_t(255); }
constexpr size_t NA_NBC = (NA_MAX_COLS + 255) / 256;
constexpr size_t NA_OFF_SC = 0;
constexpr size_t NA_OFF_P0 = NA_OFF_SC + na_align(sizeof(stp::SolverScalars));
constexpr size_t NA_OFF_P2 = NA_OFF_P0 + na_align(8 * stp::SPMV_MAX_BLOCKS);
constexpr size_t NA_OFF_P3 = NA_OFF_P2 + na_align(8 * stp::SPMV_MAX_BLOCKS);
constexpr size_t NA_OFF_P1 = NA_OFF_P3 + na_align(8 * stp::SPMV_MAX_BLOCKS);
constexpr size_t NA_OFF_P4 = NA_OFF_P1 + na_align(8 * NA_NBC);
constexpr size_t NA_OFF_P5 = NA_OFF_P4 + na_align(8 * NA_NBC);
constexpr size_t NA_OFF_R = NA_OFF_P5 + na_align(8 * NA_NBC);
constexpr size_t NA_OFF_Q = NA_OFF_R + na_align(8 * NA_MAX_ROWS);
constexpr size_t NA_OFF_P = NA_OFF_Q + na_align(4 * NA_MAX_ROWS);
constexpr size_t NA_OFF_S = NA_OFF_P + na_align(4 * NA_MAX_COLS);
constexpr size_t NA_OFF_Z = NA_OFF_S + na_align(4 * NA_MAX_COLS);
constexpr size_t NA_OFF_SP = NA_OFF_Z + na_align(4 * NA_MAX_COLS);
constexpr size_t NA_OFF_DINV = NA_OFF_SP + na_align(4 * NA_MAX_COLS);
constexpr size_t NA_OFF_V32 = NA_OFF_DINV + na_align(4 * NA_MAX_COLS);
constexpr size_t NA_OFF_Y = NA_OFF_V32 + na_align(4 * NA_MAX_ROWS);
constexpr size_t NA_OFF_G = NA_OFF_Y + na_align(8 * NA_MAX_ROWS);
constexpr size_t NA_OFF_NORM = NA_OFF_G + na_align(8 * NA_MAX_COLS);
constexpr size_t NA_BYTES = NA_OFF_NORM + na_align(8 * 2 * 256);   // 52.1 MiB
constexpr int NA_MAX_DEVICES = 64;

std::mutex g_scratch_mu;
void* g_scratch_ptr[NA_MAX_DEVICES] = {};

}  // namespace

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

// Device address of this GPU's copy of `scratch` (looked up once per device,
// with that device current).
static int scratch_address(unsigned char** out) {
    int dev = 0;
    STP_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NA_MAX_DEVICES) return fail(STP_ERR_INVALID, "stp_cgls_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "stp_cgls_noalloc: hipGetSymbolAddress(scratch)");
        g_scratch_ptr[dev] = p;
    }
    *out = static_cast<unsigned char*>(g_scratch_ptr[dev]);
    return STP_OK;
}

extern "C" {

int stp_cgls(const stp_csr* hc, const double* b, double lam, double tol, int32_t max_iter, int32_t precond,
             void* precond_handle, double* x, double* history, stp_solve_info* info, void* stream) {
    stp_csr* h = const_cast<stp_csr*>(hc);
    int rc = cgls_check(h, b, lam, max_iter, precond, precond_handle, x, "stp_cgls");
    if (rc != STP_OK) return rc;
    hipStream_t st = S(stream);
    const int rows = h->rows, cols = h->cols;
    if (h->gn[0] > 0) {
        rc = stp::build_solver_layout(h, st);
        if (rc != STP_OK) return rc;
    }
    const int W = h->gn[0] > 0 ? h->lW : h->W, WT = h->gn[0] > 0 ? h->lWT : h->WT;
    const int nbq = std::max(stp::spmv_blocks(rows, W), 1);
    const int nbs = std::max(stp::spmv_blocks(cols, WT), 1);
    const int nbc = std::max((cols + 255) / 256, 1);
    STP_HIP(grow(&h->r, &h->c_r, rows));
    STP_HIP(grow(&h->p, &h->c_p, cols));
    STP_HIP(grow(&h->q, &h->c_q, rows));
    STP_HIP(grow(&h->s, &h->c_s, cols));
    STP_HIP(grow(&h->z, &h->c_z, cols));
    STP_HIP(grow(&h->sp, &h->c_sp, cols));
    STP_HIP(grow(&h->dinv, &h->c_dinv, cols));
    STP_HIP(grow(&h->r32, &h->c_r32, rows));
    STP_HIP(grow(&h->ry, &h->c_ry, rows));
    STP_HIP(grow(&h->rg, &h->c_rg, cols));
    STP_HIP(grow(&h->pnorm, &h->c_pnorm, 2 * 256));
    STP_HIP(grow(&h->part0, &h->c_p0, nbq));
    STP_HIP(grow(&h->part1, &h->c_p1, nbc));
    STP_HIP(grow(&h->part2, &h->c_p2, nbs));
    STP_HIP(grow(&h->part3, &h->c_p3, nbs));
    STP_HIP(grow(&h->part4, &h->c_p4, nbc));
    STP_HIP(grow(&h->part5, &h->c_p5, nbc));
    STP_HIP(grow(&h->sc, &h->c_sc, 1));
    CglsWork w;
    w.r = h->r; w.p = h->p; w.q = h->q; w.s = h->s; w.z = h->z; w.sp = h->sp; w.dinv = h->dinv; w.v32 = h->r32; w.y = h->ry; w.g = h->rg; w.pnorm = h->pnorm;
    w.part0 = h->part0; w.part1 = h->part1; w.part2 = h->part2; w.part3 = h->part3; w.part4 = h->part4;
    w.part5 = h->part5; w.sc = h->sc;
    OwnedEvent ev;
    w.ev = ev.ev;
    return cgls_solve(h, b, lam, tol, max_iter, precond, precond_handle, x, history, info, st, w);
}

int stp_cgls_noalloc(const stp_csr* hc, const double* b, double lam, double tol, int32_t max_iter, int32_t precond,
                     void* precond_handle, double* x, double* history, stp_solve_info* info, void* stream) {
    stp_csr* h = const_cast<stp_csr*>(hc);
    int rc = cgls_check(h, b, lam, max_iter, precond, precond_handle, x, "stp_cgls_noalloc");
    if (rc != STP_OK) return rc;
    if (h->rows > NA_MAX_ROWS || h->cols > NA_MAX_COLS)
        return fail(STP_ERR_INVALID, "stp_cgls_noalloc: at most 1048576 rows and 1048576 columns (its static "
                                     "workspace); use stp_cgls for larger systems");
    if (h->gn[0] > 0 && !h->lay_ready)
        return fail(STP_ERR_STATE, "stp_cgls_noalloc: the matrix has a cell grid but no solver layout yet; call "
                                   "stp_csr_prepare_layout first (it allocates)");
    unsigned char* base = nullptr;
    rc = scratch_address(&base);
    if (rc != STP_OK) return rc;
    CglsWork w;
    w.sc = reinterpret_cast<stp::SolverScalars*>(base + NA_OFF_SC);
    w.part0 = reinterpret_cast<double*>(base + NA_OFF_P0);
    w.part1 = reinterpret_cast<double*>(base + NA_OFF_P1);
    w.part2 = reinterpret_cast<double*>(base + NA_OFF_P2);
    w.part3 = reinterpret_cast<double*>(base + NA_OFF_P3);
    w.part4 = reinterpret_cast<double*>(base + NA_OFF_P4);
    w.part5 = reinterpret_cast<double*>(base + NA_OFF_P5);
    w.r = reinterpret_cast<double*>(base + NA_OFF_R);
    w.q = reinterpret_cast<float*>(base + NA_OFF_Q);
    w.p = reinterpret_cast<float*>(base + NA_OFF_P);
    w.s = reinterpret_cast<float*>(base + NA_OFF_S);
    w.z = reinterpret_cast<float*>(base + NA_OFF_Z);
    w.sp = reinterpret_cast<float*>(base + NA_OFF_SP);
    w.dinv = reinterpret_cast<float*>(base + NA_OFF_DINV);
    w.v32 = reinterpret_cast<float*>(base + NA_OFF_V32);
    w.y = reinterpret_cast<double*>(base + NA_OFF_Y);
    w.g = reinterpret_cast<double*>(base + NA_OFF_G);
    w.pnorm = reinterpret_cast<double*>(base + NA_OFF_NORM);
    return cgls_solve(h, b, lam, tol, max_iter, precond, precond_handle, x, history, info, S(stream), w);
}

}  // extern "C"
