/*
 * FP64 CPU oracle for the acoustic finite-difference propagator.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the GPU path is compared
 * against and the timed CPU baseline; it is never linked into, called by, or
 * used as a fallback for the product library.
 *
 * It restates, loop for loop, the reference solver's forward stepping:
 *   - stencil pass with finite guards  : src/cpp/solver/wave_equation.cpp:299-405
 *   - sponge pass (current and previous): src/cpp/solver/wave_equation.cpp:262-297
 *     (sponge_mode 1: the Cerjan variant, next and current; not in the reference)
 *   - buffer rotation                  : src/cpp/solver/wave_equation.cpp:56-76
 *   - health check + repair            : src/cpp/solver/wave_equation.cpp:407-433,
 *                                         src/cpp/core/fault_tolerance.cpp:72-157,198-257
 *   - source injection                 : src/cpp/solver/wave_equation.cpp:39-54,118-125
 *   - forward_propagate driver         : src/cpp/solver/wave_equation.cpp:100-152
 *   - CFL check                        : src/cpp/core/fault_tolerance.cpp:49-70
 * generalised from the reference's 2nd-order 7-point Laplacian
 * (wave_equation.cpp:370-373) to a radius-R star stencil, R = order/2, with
 * the standard central-difference weights; order 2 reproduces the reference
 * arithmetic exactly.  Receiver sampling (the field after each step) is new:
 * the reference has none.
 *
 * Layout: idx = ix*nyh*nzh + iy*nzh + iz, z contiguous (src/cpp/core/grid.cpp:45-49).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "fd_oracle.h"

#ifndef ORC_REAL
#define ORC_REAL double
#endif
typedef ORC_REAL real;

enum { ORC_OK = 0, ORC_NAN_REPAIRED = 1, ORC_CFL_VIOLATION = 2, ORC_FATAL = 3 };

/* central second-derivative weights, a[0] centre, a[k] at +-k */
static const double W2[2] = {-2.0, 1.0};
static const double W8[5] = {-205.0 / 72.0, 8.0 / 5.0, -1.0 / 5.0, 8.0 / 315.0, -1.0 / 560.0};

int orc_fd_weights(int order, double* out) {
    if (order == 2) { memcpy(out, W2, sizeof W2); return 1; }
    if (order == 8) { memcpy(out, W8, sizeof W8); return 4; }
    return -1;
}

static inline int is_valid(real v) { return isfinite((double)v); }

/* reference FaultTolerance::safe_clamp (fault_tolerance.cpp:86-89) */
static inline real safe_clamp(real v, real lo, real hi) {
    if (!is_valid(v)) return 0.0;
    return v < lo ? lo : (v > hi ? hi : v);
}

static double vmax_of(const real* vel, int64_t n) {
    double vmax = 0.0;
    for (int64_t i = 0; i < n; ++i)
        if (is_valid(vel[i]) && (double)vel[i] > vmax) vmax = vel[i];
    return vmax;
}

/* max leapfrog-stable v*dt/h for a cubic cell: sqrt(4 / (3*|symbol(pi)|)) */
double orc_fd_stability_limit(int order) {
    double a[5];
    int R = orc_fd_weights(order, a);
    if (R < 0) return 0.0;
    double s = a[0];
    for (int k = 1; k <= R; ++k) s += 2.0 * a[k] * ((k & 1) ? -1.0 : 1.0);
    return sqrt(4.0 / (3.0 * fabs(s)));
}

/* verify_cfl_condition (fault_tolerance.cpp:49-70): per-axis vmax*dt/h <= limit.
 * For order 8 the leapfrog stability bound of the wider stencil is also
 * enforced (the reference's 0.49 is above it). */
int orc_fd_verify_cfl(const orc_fd_cfg* c, const real* vel) {
    int64_t n = (c->nx + 2 * c->hw) * (c->ny + 2 * c->hw) * (c->nz + 2 * c->hw);
    double vmax = vmax_of(vel, n);
    double cx = vmax * c->dt / c->dx, cy = vmax * c->dt / c->dy, cz = vmax * c->dt / c->dz;
    double cm = fmax(cx, fmax(cy, cz));
    if (!(cm <= c->cfl_safety_factor)) return 0;
    if (c->order != 2) {
        double a[5];
        int R = orc_fd_weights(c->order, a);
        double s = a[0];
        for (int k = 1; k <= R; ++k) s += 2.0 * a[k] * ((k & 1) ? -1.0 : 1.0);
        double lam = fabs(s) * (1.0 / (c->dx * c->dx) + 1.0 / (c->dy * c->dy) + 1.0 / (c->dz * c->dz));
        if (!(vmax * vmax * c->dt * c->dt * lam <= 4.0)) return 0;
    }
    return 1;
}

/* Stencil pass: reference fd_7point_operator_safe (wave_equation.cpp:299-405). */
int orc_fd_stencil(const orc_fd_cfg* c, const real* uc, const real* up, real* un, const real* vel) {
    const int64_t nxh = c->nx + 2 * c->hw, nyh = c->ny + 2 * c->hw, nzh = c->nz + 2 * c->hw;
    const int64_t sx = nyh * nzh, sy = nzh;
    const int hw = c->hw;
    double a[5];
    const int R = orc_fd_weights(c->order, a);
    if (R < 0 || R > hw) return -1;
    const double idx2 = 1.0 / (c->dx * c->dx), idy2 = 1.0 / (c->dy * c->dy), idz2 = 1.0 / (c->dz * c->dz);
    const double dt2 = c->dt * c->dt;
    const double local_vmax = vmax_of(vel, nxh * sx);
    const double cfl = fmax(local_vmax * c->dt / c->dx, fmax(local_vmax * c->dt / c->dy, local_vmax * c->dt / c->dz));
    const int cfl_violation = cfl > 1.0;
    double cfl_damp = 1.0;
    if (cfl_violation) {
        double hmin = fmin(c->dx, fmin(c->dy, c->dz));
        double ratio = local_vmax * c->dt / hmin;
        double f = c->cfl_safety_factor / ratio;
        if (f > 1.0) f = 1.0;
        cfl_damp = f * f;
    }
    const real lim = (real)c->amplitude_limit;
    int status = ORC_OK;
#pragma omp parallel for schedule(static) reduction(max : status)
    for (int64_t ix = hw; ix < nxh - hw; ++ix) {
        for (int64_t iy = hw; iy < nyh - hw; ++iy) {
            for (int64_t iz = hw; iz < nzh - hw; ++iz) {
                const int64_t i = ix * sx + iy * sy + iz;
                const real u = uc[i], p = up[i];
                if (!is_valid(u) || !is_valid(p)) { un[i] = 0.0; status = ORC_NAN_REPAIRED; continue; }
                int bad = 0;
                for (int k = 1; k <= R; ++k) {
                    bad |= !is_valid(uc[i + k * sx]) | !is_valid(uc[i - k * sx]);
                    bad |= !is_valid(uc[i + k * sy]) | !is_valid(uc[i - k * sy]);
                    bad |= !is_valid(uc[i + k]) | !is_valid(uc[i - k]);
                }
                if (bad) { un[i] = u; status = ORC_NAN_REPAIRED; continue; }
                /* reference order: x term + y term + z term, each (sum of taps)/h^2 */
                real lx = (real)a[0] * u, ly = (real)a[0] * u, lz = (real)a[0] * u;
                if (R == 1) {
                    lx = uc[i + sx] - 2.0 * u + uc[i - sx];
                    ly = uc[i + sy] - 2.0 * u + uc[i - sy];
                    lz = uc[i + 1] - 2.0 * u + uc[i - 1];
                } else {
                    for (int k = 1; k <= R; ++k) {
                        lx += (real)a[k] * (uc[i + k * sx] + uc[i - k * sx]);
                        ly += (real)a[k] * (uc[i + k * sy] + uc[i - k * sy]);
                        lz += (real)a[k] * (uc[i + k] + uc[i - k]);
                    }
                }
                const real lap = lx * (real)idx2 + ly * (real)idy2 + lz * (real)idz2;
                real v = vel[i];
                if (!is_valid(v) || v <= 0.0) v = 1500.0;
                real upd = v * v * (real)dt2 * lap;
                if (cfl_violation) upd *= (real)cfl_damp;
                real nv = 2.0 * u - p + upd;
                if (!is_valid(nv)) { nv = u; status = ORC_NAN_REPAIRED; }
                un[i] = safe_clamp(nv, -lim, lim);
            }
        }
    }
    return status;
}

/* sponge damping factor at with-halo coords; 1.0 outside the sponge.
 * Reference apply_sponge_layer (wave_equation.cpp:262-297). */
double orc_fd_sponge_factor(const orc_fd_cfg* c, int64_t ix, int64_t iy, int64_t iz) {
    const int64_t nxh = c->nx + 2 * c->hw, nyh = c->ny + 2 * c->hw, nzh = c->nz + 2 * c->hw;
    const int hw = c->hw, sw = c->sponge_width;
    int d = sw;
#define ORC_AX(i, nh)                                                   \
    if ((i) < hw + sw && (int)((i) - hw) < d) d = (int)((i) - hw);      \
    if ((i) >= (nh) - hw - sw && (int)((nh) - hw - (i) - 1) < d) d = (int)((nh) - hw - (i) - 1);
    ORC_AX(ix, nxh) ORC_AX(iy, nyh) ORC_AX(iz, nzh)
#undef ORC_AX
    if (d < sw && d >= 0) return 1.0 - c->sponge_damping * (double)(sw - d) * (double)(sw - d) / (double)(sw * sw);
    return 1.0;
}

void orc_fd_sponge(const orc_fd_cfg* c, real* cur, real* prev) {
    const int64_t nxh = c->nx + 2 * c->hw, nyh = c->ny + 2 * c->hw, nzh = c->nz + 2 * c->hw;
    if (c->sponge_width <= 0) return;
#pragma omp parallel for schedule(static)
    for (int64_t ix = 0; ix < nxh; ++ix)
        for (int64_t iy = 0; iy < nyh; ++iy)
            for (int64_t iz = 0; iz < nzh; ++iz) {
                const double f = orc_fd_sponge_factor(c, ix, iy, iz);
                if (f != 1.0) {
                    const int64_t i = (ix * nyh + iy) * nzh + iz;
                    cur[i] *= (real)f;
                    prev[i] *= (real)f;
                }
            }
}

int64_t orc_count_nan_inf(const real* d, int64_t n) {
    int64_t k = 0;
    for (int64_t i = 0; i < n; ++i) k += !is_valid(d[i]);
    return k;
}

void orc_clamp_field(real* d, int64_t n, double lo, double hi) {
    for (int64_t i = 0; i < n; ++i) d[i] = safe_clamp(d[i], (real)lo, (real)hi);
}

/* FaultTolerance::repair_field_nan (fault_tolerance.cpp:98-157); nx,ny,nz are
 * the with-halo extents.  Sequential: a repaired cell is visible to later cells. */
int64_t orc_repair_field_nan(real* d, int64_t count, int64_t nx, int64_t ny, int64_t nz, int hw, double repl) {
    int64_t repaired = 0;
    const int64_t nynz = ny * nz;
    for (int64_t ix = hw; ix < nx - hw; ++ix)
        for (int64_t iy = hw; iy < ny - hw; ++iy)
            for (int64_t iz = hw; iz < nz - hw; ++iz) {
                const int64_t i = ix * nynz + iy * nz + iz;
                if (is_valid(d[i])) continue;
                double s = 0.0;
                int n = 0;
                if (ix > hw && is_valid(d[i - nynz])) { s += d[i - nynz]; ++n; }
                if (ix < nx - hw - 1 && is_valid(d[i + nynz])) { s += d[i + nynz]; ++n; }
                if (iy > hw && is_valid(d[i - nz])) { s += d[i - nz]; ++n; }
                if (iy < ny - hw - 1 && is_valid(d[i + nz])) { s += d[i + nz]; ++n; }
                if (iz > hw && is_valid(d[i - 1])) { s += d[i - 1]; ++n; }
                if (iz < nz - hw - 1 && is_valid(d[i + 1])) { s += d[i + 1]; ++n; }
                d[i] = n > 0 ? (real)(s / n) : (real)repl;
                ++repaired;
            }
    for (int64_t i = 0; i < count; ++i)
        if (!is_valid(d[i])) { d[i] = (real)repl; ++repaired; }
    return repaired;
}

/* FaultTolerance::attempt_recovery (fault_tolerance.cpp:198-257), 3 attempts */
static int attempt_recovery(const orc_fd_cfg* c, real* cur, real* prev, int64_t n, orc_fault_stats* st) {
    const int64_t nxh = c->nx + 2 * c->hw, nyh = c->ny + 2 * c->hw, nzh = c->nz + 2 * c->hw;
    int64_t nan = orc_count_nan_inf(cur, n);
    if (nan == 0) return 1;
    st->total_nan_count += nan;
    for (int attempt = 0; attempt < 3; ++attempt) {
        orc_repair_field_nan(cur, n, nxh, nyh, nzh, c->hw, 0.0);
        orc_repair_field_nan(prev, n, nxh, nyh, nzh, c->hw, 0.0);
        orc_clamp_field(cur, n, -c->amplitude_limit, c->amplitude_limit);
        orc_clamp_field(prev, n, -c->amplitude_limit, c->amplitude_limit);
        int64_t post = orc_count_nan_inf(cur, n);
        st->cells_repaired += nan - post;
        if (post == 0) { st->timesteps_recovered++; return 1; }
        nan = post;
    }
    return 0;
}

/* One step_forward (wave_equation.cpp:56-76) on three rotating buffers.
 * bufs[0]=current, bufs[1]=previous, bufs[2]=next; rotated in place. */
int orc_fd_step(const orc_fd_cfg* c, real** bufs, const real* vel, orc_fault_stats* st) {
    const int64_t n = (c->nx + 2 * c->hw) * (c->ny + 2 * c->hw) * (c->nz + 2 * c->hw);
    int fd = orc_fd_stencil(c, bufs[0], bufs[1], bufs[2], vel);
    if (fd < 0) return fd;
    /* reference: the current and previous fields are damped before the swap,
     * so the damped current becomes the new previous and the damped previous
     * is discarded; the recurrence u^{t+1} = 2u^t - D u^{t-1} + C L u^t then
     * has the root 1 + sqrt(1 - D) > 1 for smooth modes inside the sponge
     * (it grows).  Cerjan mode damps the next and current fields instead:
     * u^{t+1} = D (2u^t - D u^{t-1} + C L u^t), whose smooth-mode root is D. */
    if (c->sponge_mode == 1) orc_fd_sponge(c, bufs[2], bufs[0]);
    else orc_fd_sponge(c, bufs[0], bufs[1]);
    real* t = bufs[1]; bufs[1] = bufs[0]; bufs[0] = t;   /* swap(previous, current) */
    t = bufs[0]; bufs[0] = bufs[2]; bufs[2] = t;         /* swap(current, next) */
    if (orc_count_nan_inf(bufs[0], n) > 0) {
        if (!attempt_recovery(c, bufs[0], bufs[1], n, st)) return ORC_FATAL;
        return ORC_NAN_REPAIRED;
    }
    return fd;
}

/* inject_source (wave_equation.cpp:39-54): interior coords, non-finite skipped */
void orc_fd_inject(const orc_fd_cfg* c, real* cur, int64_t ix, int64_t iy, int64_t iz, double w) {
    const int64_t nyh = c->ny + 2 * c->hw, nzh = c->nz + 2 * c->hw;
    const int64_t n = (c->nx + 2 * c->hw) * nyh * nzh;
    const int64_t i = (ix + c->hw) * nyh * nzh + (iy + c->hw) * nzh + (iz + c->hw);
    if (i >= 0 && i < n && isfinite(w)) cur[i] += (real)w;
}

typedef struct {
    int32_t status;
    int32_t timesteps_completed;
    int32_t completed_successfully;
    int64_t total_nan_encountered;
    int64_t total_cells_repaired;
} orc_prop_result;

/* forward_propagate (wave_equation.cpp:100-152).  cur/prev are in/out with-halo
 * arrays (zeroed on entry, as reset_field does).  src_* are interior coords of the
 * source (reference: nx/2, ny/2, nz/2).  traces[r*nt + t] = current[rcv_idx[r]]
 * after step t (receivers are new; the reference records none). */
void orc_fd_forward(const orc_fd_cfg* c, const real* vel, const double* wavelet, int32_t nw, int32_t nt,
                    int64_t src_ix, int64_t src_iy, int64_t src_iz, const int64_t* rcv_idx, int32_t nrcv,
                    real* traces, real* cur_out, real* prev_out, orc_prop_result* res) {
    const int64_t n = (c->nx + 2 * c->hw) * (c->ny + 2 * c->hw) * (c->nz + 2 * c->hw);
    real* b0 = (real*)calloc((size_t)n, sizeof(real));
    real* b1 = (real*)calloc((size_t)n, sizeof(real));
    real* b2 = (real*)calloc((size_t)n, sizeof(real));
    real* bufs[3] = {b0, b1, b2};
    orc_fault_stats st = {0, 0, 0};
    memset(res, 0, sizeof *res);
    if (!orc_fd_verify_cfl(c, vel)) {
        res->status = ORC_CFL_VIOLATION;
    } else {
        for (int32_t t = 0; t < nt; ++t) {
            if (t < nw) orc_fd_inject(c, bufs[0], src_ix, src_iy, src_iz, wavelet[t]);
            int s = orc_fd_step(c, bufs, vel, &st);
            if (s == ORC_NAN_REPAIRED) {
                res->total_nan_encountered += orc_count_nan_inf(bufs[0], n);
                res->status = ORC_NAN_REPAIRED;
            } else if (s == ORC_CFL_VIOLATION || s == ORC_FATAL || s < 0) {
                res->status = s < 0 ? ORC_FATAL : s;
                res->timesteps_completed = t;
                goto done;
            }
            for (int32_t r = 0; r < nrcv; ++r) traces[(int64_t)r * nt + t] = bufs[0][rcv_idx[r]];
            ++res->timesteps_completed;
        }
        res->completed_successfully = 1;
    }
done:
    res->total_cells_repaired = st.cells_repaired;
    memcpy(cur_out, bufs[0], (size_t)n * sizeof(real));
    memcpy(prev_out, bufs[1], (size_t)n * sizeof(real));
    free(b0); free(b1); free(b2);
}

/* Generic stepping from a given state (no reset, no CFL check): used by the
 * adjoint/RTM oracles and by parity tests that start from a non-zero field.
 * cur/prev in/out.  Injection of w[t] at each listed source cell (with-halo
 * flat index) before each step, sampling after. */
int orc_fd_run(const orc_fd_cfg* c, const real* vel, real* cur, real* prev, int32_t nsteps,
               const int64_t* src_idx, const double* src_w, int32_t nsrc, int32_t nw,
               const int64_t* rcv_idx, int32_t nrcv, real* traces) {
    const int64_t n = (c->nx + 2 * c->hw) * (c->ny + 2 * c->hw) * (c->nz + 2 * c->hw);
    real* b2 = (real*)calloc((size_t)n, sizeof(real));
    real* bufs[3] = {cur, prev, b2};
    orc_fault_stats st = {0, 0, 0};
    int agg = ORC_OK;
    for (int32_t t = 0; t < nsteps; ++t) {
        if (t < nw)
            for (int32_t s = 0; s < nsrc; ++s) {
                const double w = src_w[(int64_t)s * nw + t];
                if (isfinite(w)) bufs[0][src_idx[s]] += (real)w;
            }
        int s = orc_fd_step(c, bufs, vel, &st);
        if (s > agg) agg = s;
        for (int32_t r = 0; r < nrcv; ++r) traces[(int64_t)r * nsteps + t] = bufs[0][rcv_idx[r]];
    }
    /* the buffers rotated: copy the logical current/previous back to the caller's arrays */
    real* t0 = (real*)malloc((size_t)n * sizeof(real));
    real* t1 = (real*)malloc((size_t)n * sizeof(real));
    memcpy(t0, bufs[0], (size_t)n * sizeof(real));
    memcpy(t1, bufs[1], (size_t)n * sizeof(real));
    memcpy(cur, t0, (size_t)n * sizeof(real));
    memcpy(prev, t1, (size_t)n * sizeof(real));
    free(t0); free(t1); free(b2);
    return agg;
}

/* Ricker wavelet w(t) = (1 - 2 pi^2 f^2 tau^2) exp(-pi^2 f^2 tau^2), tau = t*dt - t0 */
void orc_ricker(double f, double dt, double t0, int32_t n, double* out) {
    const double pi = 3.14159265358979323846;
    for (int32_t i = 0; i < n; ++i) {
        const double tau = i * dt - t0, a = pi * pi * f * f * tau * tau;
        out[i] = (1.0 - 2.0 * a) * exp(-a);
    }
}
