/*
 * FP64 CPU oracle for the tomography solve.  TEST INFRASTRUCTURE ONLY: the
 * checker for the GPU sparse path and the timed CPU baseline.
 *
 * Restates:
 *   CSRMatrix::build_from_triplets  src/cpp/sparse/csr_matrix.cpp:25-72
 *       (canonical CSR before pad_to_alignment; duplicates summed in input
 *        order — the reference's std::sort is not stable, so with three or
 *        more duplicates of one (row, col) its summation order is unspecified)
 *   CSRMatrix::spmv                 csr_matrix.cpp:118-149 (even/odd accumulators)
 *   CSRMatrix::spmv_transpose       csr_matrix.cpp:151-166 (accumulator zeroed)
 *   SparseOps::conjugate_gradient   src/cpp/sparse/sparse_ops.cpp:8-50 (x0 = 0)
 *   SparseOps::jacobi_preconditioned_cg  sparse_ops.cpp:52-113 (x0 = 0)
 *   SparseOps::dot_product          sparse_ops.cpp:172-187 (4-way unrolled)
 *   SparseOps::smooth_gradient      sparse_ops.cpp:115-158 (direct 3-D sum)
 * and adds (no reference counterpart):
 *   CGLS on (A^T A + lambda I) x = A^T b, plain or Jacobi preconditioned;
 *   Siddon / Amanatides-Woo exact straight-ray path lengths through a grid of
 *   cubic cells (the matrix the reference samples matrix-free in
 *   src/python/seismic_tomography/inversion.py:329-383).
 * Compiled without FP contraction so the Siddon arithmetic is reproducible.
 */
#pragma GCC optimize("fp-contract=off")
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct { int64_t col; double val; int64_t ord; } orc_ent;

static int ent_cmp(const void* a, const void* b) {
    const orc_ent* x = (const orc_ent*)a;
    const orc_ent* y = (const orc_ent*)b;
    if (x->col != y->col) return x->col < y->col ? -1 : 1;
    return x->ord < y->ord ? -1 : (x->ord > y->ord);
}

/* returns nnz; col_out/val_out need capacity n */
int64_t orc_csr_from_triplets(int64_t rows, int64_t cols, int64_t n, const int64_t* r, const int64_t* c,
                              const double* v, int64_t* row_ptr, int64_t* col_out, double* val_out) {
    int64_t* cnt = (int64_t*)calloc((size_t)rows + 1, sizeof(int64_t));
    for (int64_t i = 0; i < n; ++i)
        if (r[i] >= 0 && r[i] < rows && c[i] >= 0 && c[i] < cols) cnt[r[i] + 1]++;
    for (int64_t i = 0; i < rows; ++i) cnt[i + 1] += cnt[i];
    const int64_t m = cnt[rows];
    orc_ent* e = (orc_ent*)malloc(sizeof(orc_ent) * (size_t)(m > 0 ? m : 1));
    int64_t* pos = (int64_t*)malloc(sizeof(int64_t) * (size_t)(rows + 1));
    memcpy(pos, cnt, sizeof(int64_t) * (size_t)(rows + 1));
    for (int64_t i = 0; i < n; ++i)
        if (r[i] >= 0 && r[i] < rows && c[i] >= 0 && c[i] < cols) {
            orc_ent* d = &e[pos[r[i]]++];
            d->col = c[i]; d->val = v[i]; d->ord = i;
        }
    int64_t nnz = 0;
    row_ptr[0] = 0;
    for (int64_t i = 0; i < rows; ++i) {
        const int64_t b = cnt[i], en = cnt[i + 1];
        qsort(e + b, (size_t)(en - b), sizeof(orc_ent), ent_cmp);
        for (int64_t k = b; k < en; ++k) {
            if (nnz > row_ptr[i] && col_out[nnz - 1] == e[k].col) val_out[nnz - 1] += e[k].val;
            else { col_out[nnz] = e[k].col; val_out[nnz] = e[k].val; ++nnz; }
        }
        row_ptr[i + 1] = nnz;
    }
    free(pos); free(e); free(cnt);
    return nnz;
}

void orc_csr_spmv(int64_t rows, const int64_t* rp, const int64_t* ci, const double* val, const double* x,
                  double* y) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < rows; ++i) {
        double s0 = 0.0, s1 = 0.0;
        int64_t p = rp[i];
        const int64_t e = rp[i + 1];
        for (; p + 1 < e; p += 2) {
            s0 += val[p] * x[ci[p]];
            s1 += val[p + 1] * x[ci[p + 1]];
        }
        for (; p < e; ++p) s0 += val[p] * x[ci[p]];
        y[i] = s0 + s1;
    }
}

void orc_csr_spmv_t(int64_t rows, int64_t cols, const int64_t* rp, const int64_t* ci, const double* val,
                    const double* x, double* y) {
    memset(y, 0, sizeof(double) * (size_t)cols);
    for (int64_t i = 0; i < rows; ++i) {
        const double xi = x[i];
        for (int64_t p = rp[i]; p < rp[i + 1]; ++p) y[ci[p]] += val[p] * xi;
    }
}

double orc_dot(const double* a, const double* b, int64_t n) {
    double s = 0.0;
    int64_t i = 0;
    for (; i + 3 < n; i += 4) s += a[i] * b[i] + a[i + 1] * b[i + 1] + a[i + 2] * b[i + 2] + a[i + 3] * b[i + 3];
    for (; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* conjugate_gradient: returns iterations run; *relres = sqrt(r.r)/|b| at exit */
int orc_cg(int64_t n, const int64_t* rp, const int64_t* ci, const double* val, const double* b, int max_iter,
           double tol, double* x, double* relres) {
    double* r = (double*)malloc(sizeof(double) * (size_t)n);
    double* p = (double*)malloc(sizeof(double) * (size_t)n);
    double* ap = (double*)malloc(sizeof(double) * (size_t)n);
    memset(x, 0, sizeof(double) * (size_t)n);
    memcpy(r, b, sizeof(double) * (size_t)n);
    memcpy(p, b, sizeof(double) * (size_t)n);
    double rs = orc_dot(r, r, n);
    const double bn = sqrt(orc_dot(b, b, n));
    int it = 0;
    *relres = bn < 1e-30 ? 0.0 : sqrt(rs) / bn;
    if (bn >= 1e-30) {
        for (it = 0; it < max_iter;) {
            orc_csr_spmv(n, rp, ci, val, p, ap);
            const double pap = orc_dot(p, ap, n);
            if (fabs(pap) < 1e-30) break;
            const double a = rs / pap;
            for (int64_t i = 0; i < n; ++i) { x[i] += a * p[i]; r[i] -= a * ap[i]; }
            const double rn = orc_dot(r, r, n);
            ++it;
            *relres = sqrt(rn) / bn;
            if (sqrt(rn) / bn < tol) break;
            const double be = rn / rs;
            for (int64_t i = 0; i < n; ++i) p[i] = r[i] + be * p[i];
            rs = rn;
        }
    }
    free(r); free(p); free(ap);
    return it;
}

/* jacobi_preconditioned_cg (sparse_ops.cpp:52-113) */
int orc_jacobi_pcg(int64_t n, const int64_t* rp, const int64_t* ci, const double* val, const double* b,
                   int max_iter, double tol, double* x, double* relres) {
    double* dinv = (double*)malloc(sizeof(double) * (size_t)n);
    double* r = (double*)malloc(sizeof(double) * (size_t)n);
    double* z = (double*)malloc(sizeof(double) * (size_t)n);
    double* p = (double*)malloc(sizeof(double) * (size_t)n);
    double* ap = (double*)malloc(sizeof(double) * (size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        dinv[i] = 1.0;
        for (int64_t k = rp[i]; k < rp[i + 1]; ++k)
            if (ci[k] == i) { if (fabs(val[k]) > 1e-30) dinv[i] = 1.0 / val[k]; break; }
    }
    memset(x, 0, sizeof(double) * (size_t)n);
    memcpy(r, b, sizeof(double) * (size_t)n);
    for (int64_t i = 0; i < n; ++i) z[i] = dinv[i] * r[i];
    memcpy(p, z, sizeof(double) * (size_t)n);
    double rz = orc_dot(r, z, n);
    const double bn = sqrt(orc_dot(b, b, n));
    int it = 0;
    *relres = bn < 1e-30 ? 0.0 : 1.0;
    if (bn >= 1e-30) {
        for (it = 0; it < max_iter;) {
            orc_csr_spmv(n, rp, ci, val, p, ap);
            const double pap = orc_dot(p, ap, n);
            if (fabs(pap) < 1e-30) break;
            const double a = rz / pap;
            for (int64_t i = 0; i < n; ++i) { x[i] += a * p[i]; r[i] -= a * ap[i]; }
            const double rn = sqrt(orc_dot(r, r, n));
            ++it;
            *relres = rn / bn;
            if (rn / bn < tol) break;
            for (int64_t i = 0; i < n; ++i) z[i] = dinv[i] * r[i];
            const double rzn = orc_dot(r, z, n);
            const double be = rzn / rz;
            for (int64_t i = 0; i < n; ++i) p[i] = z[i] + be * p[i];
            rz = rzn;
        }
    }
    free(dinv); free(r); free(z); free(p); free(ap);
    return it;
}

/* CGLS on (A^T A + lam I) x = A^T b (A: rows x cols), x0 = 0.
 * precond: 0 none, 1 Jacobi (diag = column sums of a_ij^2 + lam).
 * Stops when |A^T r - lam x| / |A^T b| < tol.  history (may be NULL) gets the
 * relative normal-equation residual after each iteration. */
int orc_cgls(int64_t rows, int64_t cols, const int64_t* rp, const int64_t* ci, const double* val, const double* b,
             double lam, double tol, int max_iter, int precond, double* x, double* relres, double* history) {
    double* r = (double*)malloc(sizeof(double) * (size_t)rows);
    double* q = (double*)malloc(sizeof(double) * (size_t)rows);
    double* s = (double*)malloc(sizeof(double) * (size_t)cols);
    double* z = (double*)malloc(sizeof(double) * (size_t)cols);
    double* p = (double*)malloc(sizeof(double) * (size_t)cols);
    double* dinv = (double*)malloc(sizeof(double) * (size_t)cols);
    for (int64_t j = 0; j < cols; ++j) dinv[j] = lam;
    for (int64_t i = 0; i < rows; ++i)
        for (int64_t k = rp[i]; k < rp[i + 1]; ++k) dinv[ci[k]] += val[k] * val[k];
    for (int64_t j = 0; j < cols; ++j) dinv[j] = precond && dinv[j] > 0 ? 1.0 / dinv[j] : 1.0;
    memset(x, 0, sizeof(double) * (size_t)cols);
    memcpy(r, b, sizeof(double) * (size_t)rows);
    orc_csr_spmv_t(rows, cols, rp, ci, val, r, s);
    const double s0 = sqrt(orc_dot(s, s, cols));
    for (int64_t j = 0; j < cols; ++j) { z[j] = dinv[j] * s[j]; p[j] = z[j]; }
    double gam = orc_dot(s, z, cols);
    int it = 0;
    *relres = s0 > 0 ? 1.0 : 0.0;
    if (s0 > 0) {
        for (it = 0; it < max_iter;) {
            orc_csr_spmv(rows, rp, ci, val, p, q);
            const double del = orc_dot(q, q, rows) + lam * orc_dot(p, p, cols);
            if (del <= 0) break;
            const double a = gam / del;
            for (int64_t j = 0; j < cols; ++j) x[j] += a * p[j];
            for (int64_t i = 0; i < rows; ++i) r[i] -= a * q[i];
            orc_csr_spmv_t(rows, cols, rp, ci, val, r, s);
            for (int64_t j = 0; j < cols; ++j) s[j] -= lam * x[j];
            const double sn = sqrt(orc_dot(s, s, cols));
            if (history) history[it] = sn / s0;
            ++it;
            *relres = sn / s0;
            if (sn / s0 < tol) break;
            for (int64_t j = 0; j < cols; ++j) z[j] = dinv[j] * s[j];
            const double gn = orc_dot(s, z, cols);
            const double be = gn / gam;
            for (int64_t j = 0; j < cols; ++j) p[j] = z[j] + be * p[j];
            gam = gn;
        }
    }
    free(r); free(q); free(s); free(z); free(p); free(dinv);
    return it;
}

/* Siddon straight-ray traversal of an n[0] x n[1] x n[2] grid of cubic cells
 * of size h with origin 0: writes the cells crossed (flat index
 * ix*ny*nz + iy*nz + iz) in traversal order and the FP64 length inside each.
 * Returns the number of cells (<= cap) or -1 when cap is too small.
 * Crossing parameters along each axis: t = (k*h - p0) / d for the planes k
 * strictly between the endpoints; the segment [t_a, t_b] lies in the cell
 * containing its midpoint. */
int64_t orc_siddon(const int64_t* n, double h, const double* p0, const double* p1, int64_t* cells, double* lens,
                   int64_t cap) {
    double d[3], len2 = 0.0;
    for (int a = 0; a < 3; ++a) { d[a] = p1[a] - p0[a]; len2 += d[a] * d[a]; }
    const double len = sqrt(len2);
    if (len == 0.0) return 0;
    /* merge the three sorted crossing sequences */
    double tnext[3];
    int64_t knext[3], kend[3], kdir[3];
    for (int a = 0; a < 3; ++a) {
        if (d[a] > 0) {
            kdir[a] = 1;
            knext[a] = (int64_t)floor(p0[a] / h) + 1;
            kend[a] = (int64_t)ceil(p1[a] / h) - 1;
        } else if (d[a] < 0) {
            kdir[a] = -1;
            knext[a] = (int64_t)ceil(p0[a] / h) - 1;
            kend[a] = (int64_t)floor(p1[a] / h) + 1;
        } else {
            kdir[a] = 0; knext[a] = 0; kend[a] = 0;
        }
        tnext[a] = d[a] != 0 ? ((double)knext[a] * h - p0[a]) / d[a] : INFINITY;
        if (d[a] != 0 && (kdir[a] > 0 ? knext[a] > kend[a] : knext[a] < kend[a])) tnext[a] = INFINITY;
    }
    int64_t m = 0;
    double t0 = 0.0;
    for (;;) {
        int a = 0;
        if (tnext[1] < tnext[a]) a = 1;
        if (tnext[2] < tnext[a]) a = 2;
        const double t1 = tnext[a] < 1.0 ? tnext[a] : 1.0;
        if (t1 > t0) {
            const double tm = 0.5 * (t0 + t1);
            int64_t ix = (int64_t)floor((p0[0] + tm * d[0]) / h);
            int64_t iy = (int64_t)floor((p0[1] + tm * d[1]) / h);
            int64_t iz = (int64_t)floor((p0[2] + tm * d[2]) / h);
            if (ix < 0) ix = 0; if (ix >= n[0]) ix = n[0] - 1;
            if (iy < 0) iy = 0; if (iy >= n[1]) iy = n[1] - 1;
            if (iz < 0) iz = 0; if (iz >= n[2]) iz = n[2] - 1;
            if (m >= cap) return -1;
            cells[m] = (ix * n[1] + iy) * n[2] + iz;
            lens[m] = (t1 - t0) * len;
            ++m;
        }
        if (t1 >= 1.0) break;
        t0 = t1;
        knext[a] += kdir[a];
        tnext[a] = (kdir[a] > 0 ? knext[a] > kend[a] : knext[a] < kend[a]) ? INFINITY
                                                                          : ((double)knext[a] * h - p0[a]) / d[a];
    }
    return m;
}

/* SparseOps::smooth_gradient (sparse_ops.cpp:115-158): for every cell, the
 * sum over the (2r+1)^3 cube (r = (int)(3 sigma)) of in-range cells of
 * exp(-(dx^2+dy^2+dz^2) / (2 sigma^2)) * g, divided by the in-range weight
 * sum; x-major, z contiguous. */
void orc_smooth_gradient(int64_t nx, int64_t ny, int64_t nz, double sigma, const double* g, double* out) {
    const int64_t r = (int64_t)(3 * sigma);
    for (int64_t ix = 0; ix < nx; ++ix)
        for (int64_t iy = 0; iy < ny; ++iy)
            for (int64_t iz = 0; iz < nz; ++iz) {
                double s = 0.0, ws = 0.0;
                for (int64_t dx = -r; dx <= r; ++dx) {
                    const int64_t jx = ix + dx;
                    if (jx < 0 || jx >= nx) continue;
                    for (int64_t dy = -r; dy <= r; ++dy) {
                        const int64_t jy = iy + dy;
                        if (jy < 0 || jy >= ny) continue;
                        for (int64_t dz = -r; dz <= r; ++dz) {
                            const int64_t jz = iz + dz;
                            if (jz < 0 || jz >= nz) continue;
                            const double w = exp(-(double)(dx * dx + dy * dy + dz * dz) / (2.0 * sigma * sigma));
                            s += w * g[(jx * ny + jy) * nz + jz];
                            ws += w;
                        }
                    }
                }
                out[(ix * ny + iy) * nz + iz] = s / ws;
            }
}
