/*
 * seismic_hip.h — C ABI of the MI355X (gfx950) seismic compute library.
 *
 * Every entry point returns an int status: 0 = success, negative = error
 * (stp_last_error() returns the message, thread-local).  No C++ exception
 * crosses this boundary.  Pointers named *_dev are device pointers owned by
 * the caller; pointers named *_host are host pointers owned by the caller.
 * `stream` is a hipStream_t passed as void* (NULL = the null stream).
 * Handles are not thread-safe: use one handle per host thread / stream.
 *
 * Wavefield layout (all FD entry points): flat with-halo arrays,
 *   idx = ix*nyh*nzh + iy*nzh + iz,   n?h = n? + 2*halo_width,
 * z contiguous, exactly the reference's CartesianGrid::local_index_with_halo
 * (src/cpp/core/grid.cpp:45-49).  Batched shots are stacked: shot s starts at
 * s*nxh*nyh*nzh.
 */
#ifndef SEISMIC_HIP_H
#define SEISMIC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Step status codes, same values as the reference's StepResult
 * (src/cpp/solver/wave_equation.h:24-29). */
enum {
    STP_STEP_OK = 0,
    STP_STEP_NAN_DETECTED_AND_REPAIRED = 1,
    STP_STEP_CFL_VIOLATION = 2,
    STP_STEP_FATAL_DIVERGENCE = 3
};

/* Error codes */
enum {
    STP_OK = 0,
    STP_ERR_INVALID = -1,   /* bad argument / shape */
    STP_ERR_HIP = -2,       /* HIP runtime error */
    STP_ERR_ALLOC = -3,     /* device or host allocation failed */
    STP_ERR_STATE = -4      /* call out of order (e.g. no receivers set) */
};

const char* stp_last_error(void);
int stp_version(void);

/* ---------------------------------------------------------------- device memory
 * Thin helpers so host C++ above this ABI never calls HIP itself. */
int stp_device_alloc(void** ptr_dev, int64_t bytes);
int stp_device_free(void* ptr_dev);
int stp_memcpy_h2d(void* dst_dev, const void* src_host, int64_t bytes, void* stream);
int stp_memcpy_d2h(void* dst_host, const void* src_dev, int64_t bytes, void* stream);
int stp_memcpy_d2d(void* dst_dev, const void* src_dev, int64_t bytes, void* stream);
int stp_memset_d(void* dst_dev, int value, int64_t bytes, void* stream);
int stp_stream_sync(void* stream);
int stp_device_count(int* n);

/* ---------------------------------------------------------------- FD propagator
 * Replaces WaveEquationSolver's stencil/sponge/rotation/injection loop
 * (src/cpp/solver/wave_equation.cpp:39-152, 262-405) for order 2 (reference
 * 7-point) and order 8 (radius-4 star stencil). */
typedef struct {
    int64_t nx, ny, nz;         /* interior cells, GridParams::n?_global (grid.h:9-20) */
    int32_t halo_width;         /* GridParams::halo_width; >= order/2 */
    int32_t order;              /* 2 or 8 */
    double dx, dy, dz;          /* GridParams::d? */
    double dt;                  /* SolverConfig::dt (config.h:9-19) */
    int32_t sponge_width;       /* SolverConfig::sponge_width */
    double sponge_damping;      /* SolverConfig::sponge_damping */
    double amplitude_limit;     /* WaveEquationSolver amplitude_limit_ (1e10) */
    double cfl_safety_factor;   /* FaultTolerance cfl_safety_factor_ (0.49) */
    int32_t nshots;             /* wavefields advanced per launch (>= 1) */
} stp_fd_config;

typedef struct stp_fd_plan stp_fd_plan;

int stp_fd_plan_create(const stp_fd_config* cfg, stp_fd_plan** out);
int stp_fd_plan_destroy(stp_fd_plan* plan);

/* Receivers: with-halo flat indices of interior cells (same set for every shot).
 * New: the reference samples no receivers. */
int stp_fd_set_receivers(stp_fd_plan* plan, const int64_t* idx_host, int32_t nrcv);
/* Sources: one with-halo flat index per shot (reference: the interior centre,
 * wave_equation.cpp:120-124). */
int stp_fd_set_sources(stp_fd_plan* plan, const int64_t* idx_host, int32_t nshots);

/* Sponge form.  STP_SPONGE_REFERENCE (default) is apply_sponge_layer
 * (wave_equation.cpp:262-297): the current and previous fields are damped
 * before the buffer swap, which leaves u^{t+1} = 2u^t - D u^{t-1} + C L u^t;
 * for smooth modes inside the sponge its recurrence has the root
 * 1 + sqrt(1 - D) > 1, so long runs grow there until the amplitude clamp.
 * STP_SPONGE_CERJAN (new; not in the reference) damps the next and current
 * fields instead, u^{t+1} = D (2u^t - D u^{t-1} + C L u^t), which decays
 * (root D); same memory traffic, one more multiply per cell.  The reverse-time
 * pass (stp_fd_adjoint) follows the form set here. */
#define STP_SPONGE_REFERENCE 0
#define STP_SPONGE_CERJAN 1
int stp_fd_set_sponge_mode(stp_fd_plan* plan, int32_t mode);

/* C = v^2 dt^2 per with-halo cell, with the reference's substitutions
 * (non-finite or <= 0 velocity -> 1500, wave_equation.cpp:375-378) and its CFL
 * sub-cycling damp factor folded in (wave_equation.cpp:316-331, 383-387).
 * vmax_host receives max finite velocity (compute_local_max_velocity,
 * wave_equation.cpp:435-443).  Synchronises `stream`. */
int stp_fd_coefficients(stp_fd_plan* plan, const float* vel_dev, float* coef_dev,
                        double* vmax_host, void* stream);

/* verify_cfl (wave_equation.cpp:445-451, fault_tolerance.cpp:49-70): per-axis
 * vmax*dt/h <= cfl_safety_factor; for order 8 also the leapfrog stability
 * bound of the wider stencil.  *ok = 1 if propagation may start. */
int stp_fd_verify_cfl(const stp_fd_plan* plan, double vmax, int32_t* ok);

/* Advance nsteps leapfrog steps in place.
 *   u_cur_dev  : current field u^t (all shots), pre-injection
 *   u_prev_dev : previous field; stored undamped unless prev_damped = 1
 *                (then it already carries the sponge factor, as the
 *                reference's `previous` buffer does)
 * Step s (t = t_offset + s) injects wavelet_dev[shot*nw + t] at the shot's
 * source when t < nw (non-finite samples skipped), then advances, then samples
 * receivers into traces_dev[(shot*nrcv + r)*trace_stride + s].  wavelet_dev
 * and traces_dev may be NULL.  step_status_dev[s] (may be NULL) is OR-ed with
 * STP_STEP_NAN_DETECTED_AND_REPAIRED when a finite guard fired in step s.
 * After an odd nsteps the two buffers have swapped roles: the new current
 * field is in u_prev_dev.  Asynchronous on `stream`. */
int stp_fd_propagate(stp_fd_plan* plan, float* u_cur_dev, float* u_prev_dev, const float* coef_dev,
                     int32_t nsteps, int32_t prev_damped, const float* wavelet_dev, int32_t nw,
                     int32_t t_offset, float* traces_dev, int32_t trace_stride,
                     int32_t* step_status_dev, void* stream);

/* Measurement kernel, no counterpart in the reference: out = a + b * c over n
 * FP32 values (n a multiple of 4, device arrays 16-byte aligned), one float4
 * per thread, with non-temporal loads and stores when nontemporal != 0.  It
 * moves the FD step's algorithmic bytes (three reads and one write per cell)
 * without the stencil: the measured streaming ceiling that bench.py reports
 * the step kernel against. */
int stp_stream_3r1w_f32(int64_t n, const float* a, const float* b, const float* c, float* out, int32_t nontemporal,
                        void* stream);

/* Sponge factor D per with-halo cell (1 outside the sponge, and in the halo),
 * apply_sponge_layer (wave_equation.cpp:262-297). */
int stp_fd_sponge_factors(const stp_fd_plan* plan, float* d_dev, void* stream);

/* WaveEquationSolver::compute_energy (wave_equation.cpp:482-490): sum of u^2
 * over the finite values of an FP32 device field of n cells (halo included),
 * accumulated in FP64.  Synchronises. */
int stp_field_energy_f32(const float* u_dev, int64_t n, double* energy_host, void* stream);
/* WaveEquationSolver::save_snapshot (wave_equation.cpp:463-480): the interior
 * (nxh-2hw) x (nyh-2hw) x (nzh-2hw) cells of a with-halo FP32 field as FP64,
 * x-major, non-finite values as 0.  Asynchronous. */
int stp_field_interior_f64(const float* u_dev, int64_t nxh, int64_t nyh, int64_t nzh, int32_t hw, double* out_dev,
                           void* stream);

/* ---------------------------------------------------------------- field health
 * GPU versions of FaultTolerance::count_nan_inf / clamp_field /
 * repair_field_nan (fault_tolerance.cpp:72-157), as bound by
 * check_field_health / clamp_field_safe / repair_field_nan (bindings.cpp:275-301). */
int stp_count_nonfinite_f64(const double* data_dev, int64_t n, int64_t* count_host, void* stream);
int stp_clamp_field_f64(double* data_dev, int64_t n, double lo, double hi, void* stream);
int stp_repair_field_nan_f64(double* data_dev, int64_t count, int64_t nx, int64_t ny, int64_t nz,
                             int32_t halo_width, double replace_value, void* stream);


/* ---------------------------------------------------------------- tomography
 * Device CSR of the straight-ray matrix A (FP32 values, int32 columns and row
 * offsets) plus an explicit A^T built on the device (deterministic order).
 * Replaces CSRMatrix's storage and SpMV (src/cpp/sparse/csr_matrix.h:17-71,
 * csr_matrix.cpp:118-166) and SparseOps' solvers (sparse_ops.cpp:8-113). */
typedef struct stp_csr stp_csr;

typedef struct {
    int32_t iterations;         /* iterations run */
    int32_t converged;          /* 1 if the (true) relative residual < tol */
    double rel_residual;        /* recursive residual at exit */
    double true_rel_residual;   /* residual recomputed in FP64 at exit */
    int32_t replacements;       /* FP64 residual replacements (refinement restarts) */
} stp_solve_info;

int stp_csr_create(int64_t rows, int64_t cols, int64_t nnz, const int32_t* row_ptr_dev, const int32_t* col_dev,
                   const float* val_dev, stp_csr** out, void* stream);
/* Exact straight-ray (Siddon) matrix: ray r runs from pts[6r..6r+2] to
 * pts[6r+3..6r+5] (device FP64) through an n[0] x n[1] x n[2] grid of cubic
 * cells of size h (origin 0); entry = path length inside the cell (FP64
 * arithmetic rounded to FP32), columns ix*n1*n2 + iy*n2 + iz sorted per row.
 * New: the reference samples rays matrix-free (inversion.py:329-383). */
int stp_csr_from_rays(int64_t nrays, const int64_t* n_host, double h, const double* pts_dev, stp_csr** out,
                      void* stream);
int stp_csr_destroy(stp_csr* A);
int stp_csr_shape(const stp_csr* A, int64_t* rows, int64_t* cols, int64_t* nnz);
int stp_csr_export(const stp_csr* A, int32_t transpose, int32_t* row_ptr_dev, int32_t* col_dev, float* val_dev,
                   void* stream);
/* y = A x (transpose = 0) or A^T x; FP32 accumulation.  Once the solver
 * layout of A is built (stp_csr_prepare_layout, or a CGLS solve of a matrix
 * with a grid) both products run on it, with x and y permuted on the way in
 * and out: the same sums up to the order of FP32 additions. */
int stp_csr_spmv(const stp_csr* A, int32_t transpose, const float* x_dev, float* y_dev, void* stream);
/* same with FP64 x, FP64 accumulation and FP64 y (values stay FP32) */
int stp_csr_spmv_f64(const stp_csr* A, int32_t transpose, const double* x_dev, double* y_dev, void* stream);

/* Declare that the columns of A are the cells of an nx x ny x nz grid
 * (column ix*ny*nz + iy*nz + iz); nx = ny = nz = 0 clears it.  stp_cgls then
 * solves in a renumbered copy of A and A^T (cells in Morton order, rows
 * sorted by the Morton code of their end cells) for gather locality, built on
 * the first solve and kept with A (one more copy of A and A^T).  Results
 * equal the plain solve up to FP32 summation order.  stp_csr_from_rays sets
 * its grid.  New: the reference has no such layout. */
int stp_csr_set_grid(stp_csr* A, int64_t nx, int64_t ny, int64_t nz);
/* Build that layout now (no-op without a grid or when built).  Synchronises. */
int stp_csr_prepare_layout(stp_csr* A, void* stream);
/* The staged form of the layout's A^T that the CGLS iteration reads: rows per
 * tile (64 or 32; 0 when no tile height keeps a tile's distinct rows within
 * the staging buffer, and the iteration gathers from global memory; -1 when
 * the layout is not built) and the total length of the tiles' lists of
 * distinct rows.  New: the reference has no such layout. */
int stp_csr_staging(const stp_csr* A, int32_t* tile_rows_host, int64_t* list_entries_host);

/* Measurement, no counterpart in the reference: the gathers and the index
 * stream of one product y = A x (transpose = 0) or A^T x (1) with FP32
 * (fp64 = 0) or FP64 x, through the matrix CGLS uses (the solver layout when
 * built, in its row and column order) and the same kernel shape, without the
 * matrix values: y holds row sums of the gathered x.  The gather cost the
 * sparse products are measured against. */
int stp_csr_gather_probe(const stp_csr* A, int32_t transpose, int32_t fp64, const void* x_dev, void* y_dev,
                         void* stream);

/* Attach the FP64 values of A (device, nnz entries in the order of the
 * values given to stp_csr_create) so that stp_cg runs in the reference's
 * all-FP64 arithmetic (CSRMatrix keeps FP64 values: csr_matrix.h:68).  The
 * FP32 copy still serves SpMV and CGLS.  Copies; val64_dev = NULL detaches. */
int stp_csr_set_values_f64(stp_csr* A, const double* val64_dev, void* stream);

/* SparseOps::conjugate_gradient / jacobi_preconditioned_cg on a square A:
 * x0 = 0, stop when |r|/|b| < tol (sparse_ops.cpp:39, 97-98), breakdown when
 * |p.Ap| < 1e-30; Jacobi = first diagonal entry with |a| > 1e-30, else 1.
 * FP64 vectors and dots; FP64 matrix values when attached
 * (stp_csr_set_values_f64), else the FP32 values.  On exit the residual
 * b - A x is recomputed with those values: info->true_rel_residual is
 * |b - A x| / |b|, info->rel_residual the recursive one.  history_dev
 * (optional) gets the recursive |r|/|b| after each iteration. */
int stp_cg(const stp_csr* A, const double* b_dev, int32_t max_iter, double tol, int32_t jacobi, double* x_dev,
           double* history_dev, stp_solve_info* info, void* stream);

/* CGLS on (A^T A + lambda I) x = A^T b, x0 = 0, stop when
 * |A^T r - lambda x| / |A^T b| < tol.  precond: 0 none, 1 Jacobi
 * (1/(sum_i a_ij^2 + lambda)), 2 learned network (precond_handle = stp_lpc*,
 * flexible CG).  FP32 SpMV, FP64 x, r and dots, FP64 residual replacement.
 * For tol >= 1e-6 the iteration forms A^T r as A^T r_ref (FP64, at the start
 * of the pass) minus A^T of the change of r since then, gathered in FP32
 * (FP64 sums); a pass that stalls or fails the FP64 residual check continues
 * from the recomputed residual with r in FP64.  With a cell grid the solve runs in
 * the solver layout, whose A^T is staged in tiles (stp_csr_prepare_layout).
 * New: the reference has CG on square systems only. */
int stp_cgls(const stp_csr* A, const double* b_dev, double lambda, double tol, int32_t max_iter, int32_t precond,
             void* precond_handle, double* x_dev, double* history_dev, stp_solve_info* info, void* stream);

/* stp_cgls without any device allocation: the same solve (same kernels, same
 * results bit for bit) with its workspace in a static device array of the
 * library (52.1 MiB per GPU, held until the process ends).  No counterpart in
 * the reference: for callers that cannot allocate (tight serving loops,
 * harnesses that forbid hipMalloc).  Limits: at most 2^20 rows and 2^20
 * columns (STP_ERR_INVALID beyond); a matrix with a cell grid must already
 * have its solver layout (stp_csr_prepare_layout; STP_ERR_STATE otherwise);
 * a learned preconditioner must be prepared.  Every call on a GPU shares that
 * GPU's workspace: calls that overlap on one GPU (other streams or threads)
 * race. */
int stp_cgls_noalloc(const stp_csr* A, const double* b_dev, double lambda, double tol, int32_t max_iter,
                     int32_t precond, void* precond_handle, double* x_dev, double* history_dev, stp_solve_info* info,
                     void* stream);

/* Canonical CSR assembly, CSRMatrix::build_from_triplets semantics
 * (csr_matrix.cpp:25-72) without its padding pass: triplets whose row or
 * column is out of range are dropped, rows are ordered by column, duplicate
 * (row, col) entries are summed in input order.  Device inputs r, c (int64),
 * v (FP64) of length n; device outputs row_ptr (rows + 1), col_out and
 * val_out (capacity n); *nnz_host = merged entry count.  Synchronises. */
int stp_csr_assemble(int64_t rows, int64_t cols, int64_t n, const int64_t* r_dev, const int64_t* c_dev,
                     const double* v_dev, int64_t* row_ptr_dev, int64_t* col_out_dev, double* val_out_dev,
                     int64_t* nnz_host, void* stream);
/* Triplets of the 7-point FD matrix (build_fd_stencil_7point,
 * csr_matrix.cpp:74-116): 7 slots per cell, absent neighbours get row -1. */
int stp_csr64_fd7_triplets(int64_t nx, int64_t ny, int64_t nz, const double* vel_dev, double dx, double dy,
                           double dz, double dt, int64_t* r_dev, int64_t* c_dev, double* v_dev, void* stream);
/* CSRMatrix::scale / add_diagonal (first entry with col == row) on FP64 CSR */
int stp_csr64_scale(int64_t nnz, double* val_dev, double alpha, void* stream);
int stp_csr64_add_diagonal(int64_t rows, const int64_t* row_ptr_dev, const int64_t* col_dev, double* val_dev,
                           double alpha, void* stream);
/* FP64 dot product (SparseOps::dot_product, sparse_ops.cpp:172-187; tree order) */
int stp_dot_f64(int64_t n, const double* a_dev, const double* b_dev, double* out_host, void* stream);
/* SparseOps::smooth_gradient (sparse_ops.cpp:115-158): Gaussian weights
 * exp(-|d|^2/(2 sigma^2)) over |d_x|,|d_y|,|d_z| <= (int)(3 sigma) (<= 64),
 * renormalised by the in-range weight sum; in/out/work are nx*ny*nz FP64
 * (x-major), work is scratch.  out must not alias in.  Asynchronous. */
int stp_smooth_gradient_f64(int64_t nx, int64_t ny, int64_t nz, double sigma, const double* in_dev, double* out_dev,
                            double* work_dev, void* stream);
/* SparseOps::gaussian_filter_3d (sparse_ops.cpp:212-234): out[ix,iy,iz] =
 * exp(-|c|^2 / (2 sigma^2)) / (sum over the grid), c = (ix, iy, iz) - n / 2.0
 * per axis; out is nx*ny*nz FP64 (x-major), work holds 1024 FP64 partial
 * sums.  Asynchronous. */
int stp_gaussian_filter_3d_f64(int64_t nx, int64_t ny, int64_t nz, double sigma, double* out_dev, double* work_dev,
                               void* stream);
/* FP64/int64 CSR -> FP32/int32 arrays for stp_csr_create */
int stp_csr64_to_f32(int64_t rows, int64_t nnz, const int64_t* row_ptr_dev, const int64_t* col_dev,
                     const double* val_dev, int32_t* row_ptr32_dev, int32_t* col32_dev, float* val32_dev,
                     void* stream);

/* ---------------------------------------------------------------- reverse-time pass
 * RTM image and adjoint-state gradient of the FD propagator (new: the
 * reference's run_rtm correlates the forward field with itself and its
 * adjoint_propagate is a time-reversed forward run, wave_equation.cpp:154-194,
 * 621-688).  One call runs the forward pass from rest with checkpoints every
 * checkpoint_interval steps (<= 0: ceil(sqrt(nt))), then the reverse pass:
 * per segment, recompute and snapshot, then step the adjoint recursion
 *   phi^k = 2 phi^{k+1} - D phi^{k+2} + C L phi^{k+1} + [C(x_r)] e_r^{k-1} delta_r
 * backwards and accumulate acc_out += phi^k * s^{k-1}, acc_ill += (s^{k-1})^2,
 * where s^t is the step-t input field u~^t (snap_lap = 0: RTM zero-lag
 * cross-correlation) or its Laplacian L u~^t (snap_lap = 1: then
 * dJ/dv = 2 acc_out / v for J = 1/2 sum (d - d_obs)^2 with e = d - d_obs).
 *   inj_traces_dev : [shot][nrcv][nt] injected at the receivers (may be NULL)
 *   inj_scale_c    : multiply the injection by C at the receiver cell
 *   acc_out_dev    : [shot][nx*ny*nz], interior-indexed; NULL (with obs_dev
 *                    given) runs the forward pass and the misfit only, with no
 *                    checkpoints and no reverse pass
 *   acc_ill_dev    : [shot][nx*ny*nz] (may be NULL)
 *   traces_out_dev : forward receiver traces [shot][nrcv][nt] (may be NULL)
 *   obs_dev        : observed traces [shot][nrcv][nt]; when given, the injected
 *                    traces are the residual e = d - obs of this call's forward
 *                    pass and *misfit_host = 1/2 sum e^2 (FP64); give either
 *                    obs_dev or inj_traces_dev
 * Allocates its checkpoint store (2*ceil(nt/K) + K + 4 fields per shot) for
 * the duration of the call.  Synchronises `stream`. */
int stp_fd_adjoint(stp_fd_plan* plan, const float* coef_dev, const float* wavelet_dev, int32_t nw, int32_t nt,
                   const float* obs_dev, const float* inj_traces_dev, int32_t inj_scale_c, int32_t snap_lap,
                   int32_t checkpoint_interval, float* acc_out_dev, float* acc_ill_dev, float* traces_out_dev,
                   double* misfit_host, void* stream);

/* RTMEngine::apply_image_condition (rtm_engine.cpp:24-57): over the interior
 * nx*ny*nz, f/a := value if finite else 0 (with-halo inputs of strides
 * ny_stride, nz_stride, halo hw); image += f*a; illum += f*f (interior-indexed). */
int stp_rtm_image_condition_f64(int64_t nx, int64_t ny, int64_t nz, int32_t hw, int64_t ny_stride, int64_t nz_stride,
                                const double* fwd_dev, const double* adj_dev, double* image_dev, double* illum_dev,
                                void* stream);
/* RTMEngine::finalize_image (rtm_engine.cpp:170-182): refl = image/illum if illum > 1e-10 else 0 */
int stp_rtm_finalize_f64(int64_t n, const double* image_dev, const double* illum_dev, double* refl_dev, void* stream);
/* y += x elementwise, FP32 into FP64 (adds a reverse-time pass's FP32 image or
 * illumination into an FP64 RTMEngine accumulator; new) */
int stp_accumulate_f32_f64(int64_t n, const float* x_dev, double* y_dev, void* stream);
/* GradientComputer.compute_adjoint_gradient (gradient.py:24-40): g = -fwd*adj*(2 v dt^2) */
int stp_adjoint_gradient_f64(int64_t n, const double* fwd_dev, const double* adj_dev, const double* vel_dev, double dt,
                             double* grad_dev, void* stream);

/* ---------------------------------------------------------------- learned preconditioner
 * New: the reference has only Jacobi-PCG (sparse_ops.cpp:52-113), which stays
 * the fallback (stp_cgls precond 1).  A 3-level 3-D convolutional network
 * (16/32/64 channels, 3x3x3, ReLU; BF16 MFMA, FP32 accumulation) reads the
 * Jacobi-scaled couplings of N = A^T A + lambda I to the 26 neighbours of each
 * cell of an n0 x n1 x n2 grid (A's columns, x-major; each dimension a
 * multiple of 4) and outputs, per cell, a log scale and the log weights of
 * the three grid edges to i + e_x, i + e_y, i + e_z.  The preconditioner is
 * M = S (I + L_w) S with s_i = exp(o0_i)/sqrt(N_ii) and L_w the weighted
 * Laplacian of the 7-point grid graph (w = exp(o)): symmetric positive
 * definite for any weights, one 7-point stencil per application.  Untrained
 * (zero output layer), M = D^-1/2 (I + L) D^-1/2.  Training differentiates
 * unrolled preconditioned CG: stp_lpc_backward_accumulate gives the output
 * gradient of <g, M r>, stp_lpc_backward takes it through the network and
 * stp_lpc_adam_step updates the weights.  A handle holds one prepared system
 * at a time. */
typedef struct stp_lpc stp_lpc;
int stp_lpc_create(uint64_t seed, stp_lpc** out);
int stp_lpc_destroy(stp_lpc* h);
int stp_lpc_num_params(int64_t* n);
/* FP32 master weights, layers e1, e2, e3, d2, d1, o, each [cout][27][cin] */
int stp_lpc_get_params(const stp_lpc* h, float* host, int64_t n);
int stp_lpc_set_params(stp_lpc* h, const float* host, int64_t n);
/* network input of system A: feat_dev (optional) [cell][32] FP32 copy of the
 * BF16 features, dsq_dev (optional) [cell] = 1/sqrt(N_ii) */
int stp_lpc_features(stp_lpc* h, const stp_csr* A, const int64_t* n_host, double lambda, float* feat_dev,
                     float* dsq_dev, void* stream);
/* features + network: prepares M for A (used by stp_cgls precond 2 and stp_lpc_apply) */
int stp_lpc_prepare(stp_lpc* h, const stp_csr* A, const int64_t* n_host, double lambda, void* stream);
/* the same, keeping the activations and workspace that training needs */
int stp_lpc_prepare_train(stp_lpc* h, const stp_csr* A, const int64_t* n_host, double lambda, void* stream);
/* coefficients of M for the prepared system, [4][cell]: s, then the weights
 * of the edges (i, i + e_x), (i, i + e_y), (i, i + e_z) (0 where i + e leaves
 * the grid) */
int stp_lpc_coefficients(const stp_lpc* h, float* coef_dev, void* stream);
/* z = M s for the prepared system */
int stp_lpc_apply(const stp_lpc* h, const float* s_dev, float* z_dev, void* stream);
/* adds d<g, M r>/d(network output) to the handle's output gradient (cleared
 * by each prepare and each backward); needs stp_lpc_prepare_train */
int stp_lpc_backward_accumulate(stp_lpc* h, const float* r_dev, const float* g_dev, void* stream);
/* scale x the accumulated output gradient -> gradient of the FP32 master
 * weights, kept in the handle for stp_lpc_adam_step and copied to grad_dev
 * (optional) [num_params] */
int stp_lpc_backward(stp_lpc* h, double scale, float* grad_dev, void* stream);
/* one Adam step (beta 0.9 / 0.999, eps 1e-8) on the last backward's gradient;
 * the prepared system must be prepared again before use */
int stp_lpc_adam_step(stp_lpc* h, double lr, void* stream);
/* single convolution layer (zero padding), BF16 input [cell][cin] (raw bits),
 * FP32 weights [cout][27][cin] rounded to BF16, FP32 output [cell][cout];
 * cin, cout multiples of 16 (cout <= 96, != 80) */
int stp_conv3d_bf16(int64_t n0, int64_t n1, int64_t n2, int32_t cin, int32_t cout, const uint16_t* in_dev,
                    const float* w_dev, int32_t relu, float* out_dev, void* stream);
/* its gradients for dout [cell][cout] (BF16): din [cell][cin] (optional),
 * dw [cout][27][cin] (optional, overwritten) */
int stp_conv3d_bf16_bwd(int64_t n0, int64_t n1, int64_t n2, int32_t cin, int32_t cout, const uint16_t* in_dev,
                        const float* w_dev, const uint16_t* dout_dev, float* din_dev, float* dw_dev, void* stream);

/* ---------------------------------------------------------------- travel-time tomography
 * The straight-ray operator of the reference's Python inversion
 * (src/python/seismic_tomography/inversion.py:329-383 _forward_model and
 * :385-433 _compute_gradient), FP64 on the device.  Pick i joins src_xyz[3i..]
 * and sta_xyz[3i..] (physical coordinates, device arrays); the model is
 * n[0] x n[1] x n[2] cells (x-major, z fastest) of size h[] from origin[]
 * (host arrays).  vp_dev holds the velocity per cell. */
/* times_dev[i] = sum over the pick's samples of ds / v (reference order and
 * FP64 arithmetic, so the same value bit for bit) */
int stp_tt_forward(int64_t npicks, const double* src_xyz_dev, const double* sta_xyz_dev, const int64_t* n,
                   const double* origin, const double* h, const double* vp_dev, double* times_dev, void* stream);
/* grad_dev[cell] = sum_i r_i (-ds / v^2) w_i over the samples in the cell
 * (overwritten; r = weighted residuals as the reference passes them, weights
 * optional = 1) */
int stp_tt_gradient(int64_t npicks, const double* src_xyz_dev, const double* sta_xyz_dev, const int64_t* n,
                    const double* origin, const double* h, const double* vp_dev, const double* weighted_residuals_dev,
                    const double* weights_dev, double* grad_dev, void* stream);

/* ---------------------------------------------------------------- gradient post-processing
 * GradientComputer of the reference's Python inversion
 * (src/python/seismic_tomography/gradient.py:77-139), FP64, x-major arrays. */
/* apply_regularization (:112-126): g -= lambda * (sum over axes of the second
 * difference of the velocity, 0 at an axis' first and last index) */
int stp_gradient_tikhonov_f64(int64_t nx, int64_t ny, int64_t nz, double lambda, const double* velocity_dev,
                              double* gradient_dev, void* stream);
/* normalize_gradient (:128-139): g /= max|g| when max|g| > 1e-30; the max
 * goes to *max_abs_host when given (then the call synchronises `stream`) */
int stp_gradient_normalize_f64(int64_t n, double* gradient_dev, double* max_abs_host, void* stream);
/* build_depth_preconditioner (:96-110): 0 above water_depth_iz, else
 * 1 / (1 + (iz - water_depth_iz) * dz / 10000) */
int stp_depth_preconditioner_f64(int64_t nx, int64_t ny, int64_t nz, int64_t water_depth_iz, double dz,
                                 double* out_dev, void* stream);
/* apply_preconditioning (:79-94): g *= weights */
int stp_gradient_precondition_f64(int64_t n, const double* weights_dev, double* gradient_dev, void* stream);

/* RTMEngine::capture_single_slice (src/cpp/rtm/rtm_engine.cpp:83-165): plane
 * `index` along axis 0/1/2 (X/Y/Z) of the with-halo forward and adjoint
 * fields (strides ny_stride*nz_stride, nz_stride) and of the image condition
 * (indexed by the two in-plane coordinates, as the reference does), into
 * dim0 x dim1 outputs (X: ny x nz, Y: nx x nz, Z: nx x ny); non-finite -> 0 */
int stp_rtm_extract_slice_f64(int32_t axis, int64_t index, int64_t nx, int64_t ny, int64_t nz, int32_t halo_width,
                              int64_t ny_stride, int64_t nz_stride, const double* fwd_dev, const double* adj_dev,
                              const double* ic_dev, double* out_fwd_dev, double* out_adj_dev, double* out_ic_dev,
                              void* stream);

/* ---------------------------------------------------------------- terrain (bathymetry) boundary
 * TerrainBoundary of the reference's solver (src/cpp/solver/boundary.cpp),
 * FP64 with-halo x-major fields, bathymetry [nx_global][ny_global] (device
 * arrays).  ext (host, 9 values): nx/ny/nz with halo, halo width, global
 * offsets x/y/z of the local block, nx_global, ny_global; geo (host, 3
 * values): water depth, dz, origin_z. */
/* generate_synthetic_bathymetry (:33-54) */
int stp_terrain_bathymetry_synthetic_f64(int64_t nx, int64_t ny, double dx, double dy, double water_depth,
                                         double gradient, double roughness, double* bathy_dev, void* stream);
/* update_velocity_mask (:150-178) */
int stp_terrain_velocity_mask_f64(const int64_t* ext, const double* geo, const double* bathy_dev, double water_vp,
                                  double sediment_vp, double* velocity_dev, void* stream);
/* build_boundary_mask (:180-204): 1 subsurface, 2 water, 0 halo */
int stp_terrain_boundary_mask(const int64_t* ext, const double* geo, const double* bathy_dev, uint8_t* mask_dev,
                              void* stream);
/* apply_free_surface_bc (:80-92) */
int stp_terrain_free_surface_f64(int64_t nxh, int64_t nyh, int64_t nzh, int64_t hw, double* field_dev, void* stream);
/* apply_seafloor_bc (:94-114); field_size bounds the flat index as the
 * reference's field.size() does */
int stp_terrain_seafloor_f64(const int64_t* ext, const double* geo, const double* bathy_dev, double* field_dev,
                             int64_t field_size, void* stream);
/* apply_absorbing_bc (:116-148) */
int stp_terrain_absorbing_f64(int64_t nxh, int64_t nyh, int64_t nzh, int64_t hw, int32_t sponge_width, double damping,
                              double* field_dev, void* stream);

#ifdef __cplusplus
}
#endif
#endif
