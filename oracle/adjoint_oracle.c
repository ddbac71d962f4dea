/*
 * FP64 oracle of the reverse-time pass (RTM image and adjoint-state gradient).
 * TEST INFRASTRUCTURE ONLY.
 *
 * The reference has no discrete adjoint (its adjoint_propagate re-runs the
 * forward stepping in reversed order, src/cpp/solver/wave_equation.cpp:154-194,
 * and run_rtm correlates the forward field with itself, :621-688).  This file
 * states the adjoint of the forward recurrence restated in fd_oracle.c
 *   u~^{t+1} = 2 u~^t - D u~^{t-1} + C L u~^t + s^{t+1},  d_r^t = (u~^{t+1} - s^{t+1})(x_r)
 * for J = 1/2 sum_{t,r} (d_r^t - o_r^t)^2, with phi = C lambda:
 *   phi^k = 2 phi^{k+1} - D phi^{k+2} + C L phi^{k+1} + C(x_r) e_r^{k-1}   (k = nt..1)
 *   dJ/dC = sum_k (phi^k / C) L u~^{k-1}
 * In Cerjan sponge mode the forward step is the reference step times D, so
 * the adjoint step is too, and the receiver term gains the factor D(x_r):
 *   phi^k = D (2 phi^{k+1} - D phi^{k+2} + C L phi^{k+1}) + C(x_r) D(x_r) e_r^{k-1}
 * with the same accumulation.
 * It keeps every forward state in memory (small grids only).  Its gradient is
 * pinned in the tests against central finite differences of J computed with
 * the forward oracle.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "fd_oracle.h"

int orc_fd_step(const orc_fd_cfg* c, double** bufs, const double* vel, orc_fault_stats* st);

/* L u at interior cells (zero in the halo), same weights and h^2 as the stencil */
void orc_fd_laplacian(const orc_fd_cfg* c, const double* u, double* lap) {
    const int64_t nxh = c->nx + 2 * c->hw, nyh = c->ny + 2 * c->hw, nzh = c->nz + 2 * c->hw;
    const int64_t sx = nyh * nzh, sy = nzh;
    double a[5];
    const int R = orc_fd_weights(c->order, a);
    const double ix2 = 1.0 / (c->dx * c->dx), iy2 = 1.0 / (c->dy * c->dy), iz2 = 1.0 / (c->dz * c->dz);
    memset(lap, 0, sizeof(double) * (size_t)(nxh * sx));
#pragma omp parallel for schedule(static)
    for (int64_t ix = c->hw; ix < nxh - c->hw; ++ix)
        for (int64_t iy = c->hw; iy < nyh - c->hw; ++iy)
            for (int64_t iz = c->hw; iz < nzh - c->hw; ++iz) {
                const int64_t i = ix * sx + iy * sy + iz;
                double lx = a[0] * u[i], ly = a[0] * u[i], lz = a[0] * u[i];
                for (int k = 1; k <= R; ++k) {
                    lx += a[k] * (u[i + k * sx] + u[i - k * sx]);
                    ly += a[k] * (u[i + k * sy] + u[i - k * sy]);
                    lz += a[k] * (u[i + k] + u[i - k]);
                }
                lap[i] = lx * ix2 + ly * iy2 + lz * iz2;
            }
}

/* acc/ill: interior-indexed (nx*ny*nz).  inj: nrcv x nt (may be NULL).
 * traces_out: nrcv x nt (may be NULL).  wavelet injected at src_idx (with-halo). */
int orc_fd_adjoint(const orc_fd_cfg* c, const double* vel, const double* wavelet, int32_t nw, int32_t nt,
                   int64_t src_idx, const int64_t* rcv, int32_t nrcv, const double* inj, int32_t inj_scale_c,
                   int32_t snap_lap, double* acc, double* ill, double* traces_out) {
    const int64_t nxh = c->nx + 2 * c->hw, nyh = c->ny + 2 * c->hw, nzh = c->nz + 2 * c->hw;
    const int64_t n = nxh * nyh * nzh, nint = c->nx * c->ny * c->nz;
    double* states = (double*)malloc(sizeof(double) * (size_t)(n * nt));
    double* b0 = (double*)calloc((size_t)n, sizeof(double));
    double* b1 = (double*)calloc((size_t)n, sizeof(double));
    double* b2 = (double*)calloc((size_t)n, sizeof(double));
    double* bufs[3] = {b0, b1, b2};
    orc_fault_stats st = {0, 0, 0};
    if (!states) return -1;
    /* forward (reference-faithful stepping), keeping u~^t */
    for (int32_t t = 0; t < nt; ++t) {
        if (t < nw && isfinite(wavelet[t])) bufs[0][src_idx] += wavelet[t];
        memcpy(states + (size_t)t * n, bufs[0], sizeof(double) * (size_t)n);
        orc_fd_step(c, bufs, vel, &st);
        if (traces_out)
            for (int32_t r = 0; r < nrcv; ++r) traces_out[(int64_t)r * nt + t] = bufs[0][rcv[r]];
    }
    /* coefficients and sponge factors */
    double* C = (double*)malloc(sizeof(double) * (size_t)n);
    double* D = (double*)malloc(sizeof(double) * (size_t)n);
    for (int64_t ix = 0; ix < nxh; ++ix)
        for (int64_t iy = 0; iy < nyh; ++iy)
            for (int64_t iz = 0; iz < nzh; ++iz) {
                const int64_t i = (ix * nyh + iy) * nzh + iz;
                double v = vel[i];
                if (!isfinite(v) || v <= 0.0) v = 1500.0;
                C[i] = v * v * c->dt * c->dt;
                D[i] = orc_fd_sponge_factor(c, ix, iy, iz);
            }
    /* backward */
    double* p1 = (double*)calloc((size_t)n, sizeof(double)); /* phi^{k+1} */
    double* p2 = (double*)calloc((size_t)n, sizeof(double)); /* phi^{k+2} */
    double* pk = (double*)calloc((size_t)n, sizeof(double));
    double* lp = (double*)malloc(sizeof(double) * (size_t)n);
    double* snap = (double*)malloc(sizeof(double) * (size_t)n);
    memset(acc, 0, sizeof(double) * (size_t)nint);
    if (ill) memset(ill, 0, sizeof(double) * (size_t)nint);
    for (int32_t t = nt - 1; t >= 0; --t) {
        orc_fd_laplacian(c, p1, lp);
        for (int64_t ix = c->hw; ix < nxh - c->hw; ++ix)
            for (int64_t iy = c->hw; iy < nyh - c->hw; ++iy)
                for (int64_t iz = c->hw; iz < nzh - c->hw; ++iz) {
                    const int64_t i = (ix * nyh + iy) * nzh + iz;
                    const double f = c->sponge_mode == 1 ? D[i] : 1.0;
                    pk[i] = f * (2.0 * p1[i] - D[i] * p2[i] + C[i] * lp[i]);
                }
        if (inj)
            for (int32_t r = 0; r < nrcv; ++r) {
                const double e = inj[(int64_t)r * nt + t];
                const double f = c->sponge_mode == 1 ? D[rcv[r]] : 1.0;
                pk[rcv[r]] += f * (inj_scale_c ? C[rcv[r]] * e : e);
            }
        const double* s = states + (size_t)t * n;
        if (snap_lap) { orc_fd_laplacian(c, s, snap); s = snap; }
        for (int64_t ix = 0; ix < c->nx; ++ix)
            for (int64_t iy = 0; iy < c->ny; ++iy)
                for (int64_t iz = 0; iz < c->nz; ++iz) {
                    const int64_t i = ((ix + c->hw) * nyh + iy + c->hw) * nzh + iz + c->hw;
                    const int64_t j = (ix * c->ny + iy) * c->nz + iz;
                    acc[j] += pk[i] * s[i];
                    if (ill) ill[j] += s[i] * s[i];
                }
        double* tmp = p2; p2 = p1; p1 = pk; pk = tmp;
    }
    free(states); free(b0); free(b1); free(b2); free(C); free(D); free(p1); free(p2); free(pk); free(lp); free(snap);
    return 0;
}

/* J = 1/2 sum (d - o)^2 of a forward run (for finite-difference checks) */
double orc_fd_misfit(const orc_fd_cfg* c, const double* vel, const double* wavelet, int32_t nw, int32_t nt,
                     int64_t src_idx, const int64_t* rcv, int32_t nrcv, const double* obs) {
    const int64_t n = (c->nx + 2 * c->hw) * (c->ny + 2 * c->hw) * (c->nz + 2 * c->hw);
    double* b0 = (double*)calloc((size_t)n, sizeof(double));
    double* b1 = (double*)calloc((size_t)n, sizeof(double));
    double* b2 = (double*)calloc((size_t)n, sizeof(double));
    double* bufs[3] = {b0, b1, b2};
    orc_fault_stats st = {0, 0, 0};
    double J = 0.0;
    for (int32_t t = 0; t < nt; ++t) {
        if (t < nw && isfinite(wavelet[t])) bufs[0][src_idx] += wavelet[t];
        orc_fd_step(c, bufs, vel, &st);
        for (int32_t r = 0; r < nrcv; ++r) {
            const double e = bufs[0][rcv[r]] - obs[(int64_t)r * nt + t];
            J += 0.5 * e * e;
        }
    }
    free(b0); free(b1); free(b2);
    return J;
}
