// Learned preconditioner for CGLS on the ray-tomography normal equations
// N = A^T A + lam I (declarations: include/seismic_hip.h).
//
// The network reads the system, not the residual.  For each cell it gets the
// Jacobi-scaled couplings of N to its 26 neighbours,
//   J[i][i+d] = N[i][i+d] / sqrt(N[i][i] N[i+d][i+d]),
// plus a constant channel (which marks the model boundary through the zero
// padding), and it outputs, per cell, a log scale o0 and the log weights
// o1..o3 of the three grid edges (i, i+x), (i, i+y), (i, i+z).  With
//   s_i = exp(o0_i) / sqrt(N_ii),   w_e = W0 exp(o_e)  (edges inside the grid),
// the preconditioner is
//   M = S (I + L_w) S,   S = diag(s),  L_w = sum_e w_e (e_i - e_j)(e_i - e_j)^T,
// a weighted graph Laplacian on the 7-point grid graph: symmetric positive
// definite for any weights (I + L_w >= I), so the CG theory holds, and one
// application is a single 7-point stencil.  Untrained (zero output layer) it
// is D^-1/2 (I + W0 L) D^-1/2 with L the grid Laplacian: Jacobi scaling
// followed by a high-pass that lifts the poorly constrained short-wavelength
// components (the normal operator of straight rays damps wavenumber k like
// 1/|k|) towards the rest of the spectrum.
//
// Training differentiates the CG iteration itself: the library gives the
// gradient of <g, M r> with respect to the network output for any pair of
// vectors (stp_lpc_backward_accumulate), the host side unrolls preconditioned
// CG on training systems and back-propagates the energy-norm errors of its
// iterates, and stp_lpc_backward / stp_lpc_adam_step take that output
// gradient through the network.
//
// Network: 3 levels, 16/32/64 channels, 3x3x3 convolutions, ReLU, average
// pooling down, nearest upsampling and skip concatenation up (conv3d.hip),
// BF16 weights and activations, FP32 accumulation and FP32 master weights
// (Adam).  Layers, [cout][27][cin]:
//   e1 32->16, e2 16->32 (after pool), e3 32->64 (after pool),
//   d2 (64+32)->32, d1 (32+16)->16, o 16->16 (4 used, no ReLU).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "capi_common.h"
#include "conv3d.h"
#include "csr_handle.h"
#include "learned_precond.h"
#include "seismic_hip.h"

using stp::bf16_t;
using stp::fail;
using stp::S;

namespace {

constexpr int NFEAT = 32;   // input channels (26 couplings + constant + padding)
constexpr int NTAP = 27;    // 3x3x3 couplings and convolution taps
constexpr int NOUT = 16;    // network output channels (NCOEF used)
constexpr int NCOEF = 4;    // per cell: scale s and the weights of the +x, +y, +z edges
constexpr float W0 = 1.0f;  // edge weight of an untrained network
constexpr float OCLAMP = 30.f;  // outputs are clamped to +-OCLAMP before the exponential

struct LayerDef {
    int cin, cout;
};
constexpr LayerDef LAYERS[6] = {{32, 16}, {16, 32}, {32, 64}, {96, 32}, {48, 16}, {16, NOUT}};
enum { E1 = 0, E2, E3, D2, D1, LO };

int64_t layer_size(int l) { return (int64_t)LAYERS[l].cin * LAYERS[l].cout * NTAP; }
int64_t layer_off(int l) {
    int64_t o = 0;
    for (int k = 0; k < l; ++k) o += layer_size(k);
    return o;
}
int64_t total_params() { return layer_off(6); }

unsigned gridn(int64_t n, int bs = 256) {
    int64_t b = (n + bs - 1) / bs;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(b, 65536));
}

__device__ __forceinline__ float bf2f(bf16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ bf16_t f2bf(float f) {
    __bf16 b = (__bf16)f;
    return __builtin_bit_cast(bf16_t, b);
}

// ------------------------------------------------------------------ features
// N_ij = sum_r a_ri a_rj for each cell i and its 26 grid neighbours j, from
// the rows of A^T (cell -> sorted ray ids and lengths): N_ij is the sparse
// dot product of rows i and j.  One wavefront per cell i (four per block):
// the ray ids of row i are staged in LDS (rows longer than ROWCAP are
// searched in global memory instead); the diagonal is the lane-strided sum of
// squares of row i; for each of the 13 forward neighbours j = i + d the lanes
// take the entries of row j and look each ray up in row i by binary search,
// then the wavefront sums the products.  The value goes to channel s of i and
// channel 26 - s of j: each location has exactly one writer, no atomics, and
// the summation order is fixed.
constexpr int ROWCAP = 1024;

__device__ __forceinline__ int lower_bound_i32(const int* a, int n, int key) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int m = (lo + hi) >> 1;
        if (a[m] < key) lo = m + 1; else hi = m;
    }
    return lo;
}

__global__ void __launch_bounds__(256) couplings_kernel(int n0, int n1, int n2, const int* __restrict__ trp,
                                                        const int* __restrict__ tci,
                                                        const float* __restrict__ tval, double* nloc) {
    __shared__ int rowi[4][ROWCAP];
    const int64_t V = (int64_t)n0 * n1 * n2;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t i = (int64_t)blockIdx.x * 4 + wv;
    const bool live = i < V;
    const int a0 = live ? trp[i] : 0, a1 = live ? trp[i + 1] : 0;
    const int na = a1 - a0;
    const bool staged = na <= ROWCAP;
    if (staged)
        for (int k = lane; k < na; k += 64) rowi[wv][k] = tci[a0 + k];
    __syncthreads();
    if (!live) return;
    const int* ri = staged ? rowi[wv] : tci + a0;
    double d = 0.0;
    for (int k = a0 + lane; k < a1; k += 64) d += (double)tval[k] * (double)tval[k];
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) d += __shfl_xor(d, m);
    if (lane == 0) nloc[13 * V + i] = d;
    const int z = (int)(i % n2), y = (int)((i / n2) % n1), x = (int)(i / ((int64_t)n1 * n2));
    for (int sidx = 14; sidx < 27; ++sidx) {
        const int dx = sidx / 9 - 1, dy = (sidx / 3) % 3 - 1, dz = sidx % 3 - 1;
        if (x + dx >= n0 || y + dy < 0 || y + dy >= n1 || z + dz < 0 || z + dz >= n2) continue;
        const int64_t j = ((int64_t)(x + dx) * n1 + (y + dy)) * n2 + (z + dz);
        const int b0 = trp[j], b1 = trp[j + 1];
        double s = 0.0;
        for (int k = b0 + lane; k < b1; k += 64) {
            const int rb = tci[k];
            const int lo = lower_bound_i32(ri, na, rb);
            if (lo < na && ri[lo] == rb) s += (double)tval[a0 + lo] * (double)tval[k];
        }
#pragma unroll
        for (int m = 32; m > 0; m >>= 1) s += __shfl_xor(s, m);
        if (lane == 0) {
            nloc[(int64_t)sidx * V + i] = s;
            nloc[(int64_t)(26 - sidx) * V + j] = s;
        }
    }
}

// channels-last BF16 features and dsq = 1/sqrt(N_ii)
__global__ void features_kernel(int n0, int n1, int n2, const double* nloc, double lam, bf16_t* feat, float* dsq,
                                float* feat32) {
    const int64_t V = (int64_t)n0 * n1 * n2;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < V; i += (int64_t)gridDim.x * blockDim.x) {
        const int z = (int)(i % n2), y = (int)((i / n2) % n1), x = (int)(i / ((int64_t)n1 * n2));
        const double di = nloc[13 * V + i] + lam;
        const double si = 1.0 / sqrt(di);
        dsq[i] = (float)si;
        int ch = 0;
        for (int s = 0; s < NTAP; ++s) {
            if (s == 13) continue;
            const int dx = s / 9 - 1, dy = (s / 3) % 3 - 1, dz = s % 3 - 1;
            const int xx = x + dx, yy = y + dy, zz = z + dz;
            float f = 0.f;
            if ((unsigned)xx < (unsigned)n0 && (unsigned)yy < (unsigned)n1 && (unsigned)zz < (unsigned)n2) {
                const int64_t j = ((int64_t)xx * n1 + yy) * n2 + zz;
                const double dj = nloc[13 * V + j] + lam;
                f = (float)(nloc[s * V + i] * si / sqrt(dj));
            }
            feat[i * NFEAT + ch] = f2bf(f);
            if (feat32) feat32[i * NFEAT + ch] = bf2f(f2bf(f));
            ++ch;
        }
        for (; ch < NFEAT; ++ch) {
            const float f = ch == 26 ? 1.f : 0.f;
            feat[i * NFEAT + ch] = f2bf(f);
            if (feat32) feat32[i * NFEAT + ch] = f;
        }
    }
}

// ------------------------------------------------------------------ M = S (I + L_w) S
// network output (SoA, NOUT rows of V) -> coef[4][V]: s_i = dsq_i exp(o0_i),
// w[d]_i = W0 exp(o_{1+d}_i) for the edge (i, i + e_d) when it lies inside the
// grid, else 0 (so the stencils need no separate edge test)
__global__ void finalize_coef_kernel(int n0, int n1, int n2, const float* __restrict__ out,
                                     const float* __restrict__ dsq, float* __restrict__ coef) {
    const int64_t V = (int64_t)n0 * n1 * n2;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < V; i += (int64_t)gridDim.x * blockDim.x) {
        const int z = (int)(i % n2), y = (int)((i / n2) % n1), x = (int)(i / ((int64_t)n1 * n2));
        coef[i] = dsq[i] * expf(fminf(fmaxf(out[i], -OCLAMP), OCLAMP));
        const bool in[3] = {x + 1 < n0, y + 1 < n1, z + 1 < n2};
#pragma unroll
        for (int d = 0; d < 3; ++d)
            coef[(1 + d) * V + i] = in[d] ? W0 * expf(fminf(fmaxf(out[(1 + d) * V + i], -OCLAMP), OCLAMP)) : 0.f;
    }
}

// The 7-point graph stencil at cell i of grid (n0, n1, n2): for each axis d
// the edge to i + e_d (weight w[d]_i, zero at the far face) and the edge to
// i - e_d (weight w[d]_{i - e_d}, absent at the near face).
struct Cell {
    int64_t i, V;
    int64_t st[3];
    bool lo[3], hi[3];
};
__device__ __forceinline__ Cell cell_at(int64_t i, int n0, int n1, int n2) {
    Cell c;
    c.i = i;
    c.V = (int64_t)n0 * n1 * n2;
    const int z = (int)(i % n2), y = (int)((i / n2) % n1), x = (int)(i / ((int64_t)n1 * n2));
    c.st[0] = (int64_t)n1 * n2;
    c.st[1] = n2;
    c.st[2] = 1;
    c.lo[0] = x > 0;
    c.lo[1] = y > 0;
    c.lo[2] = z > 0;
    c.hi[0] = x + 1 < n0;
    c.hi[1] = y + 1 < n1;
    c.hi[2] = z + 1 < n2;
    return c;
}

// ((I + L_w) (s .* v))_i for the cell c; with PERM the value of cell j is
// v[perm[j]].  Every operand is loaded before the first term is formed (a
// missing neighbour reads cell i and its term is skipped), so a thread waits
// for its loads once instead of once per edge; the terms are added in the
// order +x, -x, +y, -y, +z, -z.
template <bool PERM = false>
__device__ __forceinline__ float graph_apply(const Cell& c, const float* __restrict__ coef,
                                             const float* __restrict__ v, const int* __restrict__ perm = nullptr,
                                             float* vi_out = nullptr) {
    int64_t jh[3], jl[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        jh[d] = c.hi[d] ? c.i + c.st[d] : c.i;
        jl[d] = c.lo[d] ? c.i - c.st[d] : c.i;
    }
    int64_t ki = c.i, kh[3], kl[3];   // positions of the values in v
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        kh[d] = jh[d];
        kl[d] = jl[d];
    }
    if constexpr (PERM) {
        ki = perm[c.i];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            kh[d] = perm[jh[d]];
            kl[d] = perm[jl[d]];
        }
    }
    float wh[3], wl[3], sh[3], sl[3], vh[3], vl[3];
    const float si = coef[c.i], vi = v[ki];
    if (vi_out) *vi_out = vi;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        wh[d] = coef[(1 + d) * c.V + c.i];
        wl[d] = coef[(1 + d) * c.V + jl[d]];
        sh[d] = coef[jh[d]];
        sl[d] = coef[jl[d]];
        vh[d] = v[kh[d]];
        vl[d] = v[kl[d]];
    }
    const float yi = si * vi;
    float u = yi;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        // The differences are formed for both edges whether or not they
        // exist (the empty asm keeps the compiler from moving their loads
        // into a branch), then added where the edge exists.
        float th = yi - sh[d] * vh[d], tl = yi - sl[d] * vl[d];
        asm volatile("" : "+v"(th), "+v"(tl));
        if (wh[d] != 0.f) u += wh[d] * th;   // w is zero at the far face
        if (c.lo[d]) u += wl[d] * tl;
    }
    return u;
}

// z = M r, with DOTS also the two FP64 dot partials of flexible CGLS (z.r and
// z.(r - r_prev), one per block of 256 cells, like flex_dots_kernel).  With
// PERM, r, zo and r_prev hold cell i's value at perm[i].
template <bool PERM, bool DOTS>
__global__ void __launch_bounds__(256) apply_m_kernel(int n0, int n1, int n2, const float* __restrict__ coef,
                                                      const float* __restrict__ r, float* __restrict__ zo,
                                                      const int* done, const float* __restrict__ r_prev,
                                                      double* part_zr, double* part_fl, const int* __restrict__ perm) {
    if (done && *done) return;
    const int64_t V = (int64_t)n0 * n1 * n2;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double a = 0.0, b = 0.0;
    if (i < V) {
        const Cell c = cell_at(i, n0, n1, n2);
        const int64_t pi = PERM ? (int64_t)perm[i] : i;
        float ri;   // r at cell i, from the stencil's own load
        const float z = coef[i] * graph_apply<PERM>(c, coef, r, perm, &ri);
        zo[pi] = z;
        if constexpr (DOTS) {
            const float rq = r_prev[pi];
            a = (double)z * (double)ri;
            b = (double)z * ((double)ri - (double)rq);
        }
    }
    if constexpr (!DOTS) return;
    __shared__ double red[2][256];
    red[0][threadIdx.x] = a;
    red[1][threadIdx.x] = b;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            red[0][threadIdx.x] += red[0][threadIdx.x + o];
            red[1][threadIdx.x] += red[1][threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part_zr[blockIdx.x] = red[0][0];
        part_fl[blockIdx.x] = red[1][0];
    }
}

// dacc += d<g, M r>/d(s, w).  With y = s .* r, h = s .* g, P = I + L_w:
//   d/ds_i    = g_i (P y)_i + (P h)_i r_i,
//   d/dw[d]_i = (h_i - h_j)(y_i - y_j),  j = i + e_d.
__global__ void __launch_bounds__(256) apply_m_grad_kernel(int n0, int n1, int n2, const float* __restrict__ coef,
                                                           const float* __restrict__ r,
                                                           const float* __restrict__ g, float* __restrict__ dacc) {
    const int64_t V = (int64_t)n0 * n1 * n2;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < V; i += (int64_t)gridDim.x * blockDim.x) {
        const Cell c = cell_at(i, n0, n1, n2);
        const float si = coef[i];
        dacc[i] += g[i] * graph_apply(c, coef, r) + graph_apply(c, coef, g) * r[i];
        const float yi = si * r[i], hi = si * g[i];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            if (coef[(1 + d) * V + i] == 0.f) continue;
            const int64_t j = i + c.st[d];
            dacc[(1 + d) * V + i] += (hi - coef[j] * g[j]) * (yi - coef[j] * r[j]);
        }
    }
}

// gradient with respect to the network output (channels-last BF16, NOUT
// channels): chain rule through the exponentials (zero where the output was
// clamped), channels >= NCOEF get 0
__global__ void out_grad_kernel(int64_t V, const float* __restrict__ dacc, const float* __restrict__ coef,
                                const float* __restrict__ out, float scale, bf16_t* __restrict__ d) {
    const int64_t n = V * NOUT;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const int ch = (int)(k % NOUT);
        const int64_t p = k / NOUT;
        float v = 0.f;
        if (ch < NCOEF && fabsf(out[ch * V + p]) < OCLAMP) v = scale * dacc[ch * V + p] * coef[ch * V + p];
        d[k] = f2bf(v);
    }
}

uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

template <typename T>
hipError_t dalloc(T** p, size_t n) {
    return hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * std::max<size_t>(n, 1));
}

}  // namespace

struct stp_lpc {
    // weights
    float* w32 = nullptr;
    bf16_t* wbf = nullptr;
    bf16_t* wflip = nullptr;
    float *grad = nullptr, *m1 = nullptr, *m2 = nullptr;
    float* wdw = nullptr;   // per-chunk partials of the weight gradient (CONV_BWD_WORK_MAX)
    int64_t steps = 0;
    // grid-sized workspace
    int n0 = 0, n1 = 0, n2 = 0;
    int64_t V = 0;
    bool train_ws = false;
    bf16_t *feat = nullptr, *a = nullptr, *pa = nullptr, *b = nullptr, *pb = nullptr, *c = nullptr, *cat2 = nullptr,
           *d = nullptr, *cat1 = nullptr, *e = nullptr;
    float *out = nullptr, *coef = nullptr, *dsq = nullptr;
    double* nloc = nullptr;
    // training workspace
    float *dacc = nullptr, *de = nullptr, *dcat1 = nullptr, *dd = nullptr, *da = nullptr, *dcat2 = nullptr,
          *dc = nullptr, *db = nullptr, *dpb = nullptr, *dpa = nullptr;
    bf16_t *dO = nullptr, *dpre = nullptr;
    // state of the prepared preconditioner
    bool prepared = false;
    bool dacc_live = false;   // dacc holds contributions since the last prepare / backward
};

namespace {

void free_grid(stp_lpc* h) {
    void* ps[] = {h->feat, h->a,    h->pa,   h->b,   h->pb,    h->c,   h->cat2, h->d,   h->cat1,
                  h->e,    h->out,  h->coef, h->dsq, h->nloc,  h->dacc, h->de,   h->dcat1, h->dd,
                  h->da,   h->dcat2, h->dc,  h->db,  h->dpb,   h->dpa,  h->dO,   h->dpre};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    h->feat = h->a = h->pa = h->b = h->pb = h->c = h->cat2 = h->d = h->cat1 = h->e = nullptr;
    h->out = h->coef = h->dsq = nullptr;
    h->nloc = nullptr;
    h->dacc = h->de = h->dcat1 = h->dd = h->da = h->dcat2 = h->dc = h->db = h->dpb = h->dpa = nullptr;
    h->dO = h->dpre = nullptr;
    h->V = 0;
    h->n0 = h->n1 = h->n2 = 0;
    h->train_ws = false;
    h->prepared = false;
    h->dacc_live = false;
}

void free_all(stp_lpc* h) {
    free_grid(h);
    void* ps[] = {h->w32, h->wbf, h->wflip, h->grad, h->m1, h->m2, h->wdw};
    for (void* p : ps)
        if (p) (void)hipFree(p);
}

int ensure_grid(stp_lpc* h, const int64_t* n, bool train) {
    if (n[0] < 4 || n[1] < 4 || n[2] < 4 || (n[0] | n[1] | n[2]) & 3)
        return fail(STP_ERR_INVALID, "learned preconditioner: grid dimensions must be multiples of 4 (two poolings)");
    if (n[0] * n[1] * n[2] > (int64_t)1 << 28)
        return fail(STP_ERR_INVALID, "learned preconditioner: grid too large");
    if (h->n0 == n[0] && h->n1 == n[1] && h->n2 == n[2] && (h->train_ws || !train)) return STP_OK;
    const bool keep_train = train || (h->train_ws && h->n0 == n[0] && h->n1 == n[1] && h->n2 == n[2]);
    free_grid(h);
    h->n0 = (int)n[0];
    h->n1 = (int)n[1];
    h->n2 = (int)n[2];
    const int64_t V = n[0] * n[1] * n[2], V2 = V / 8, V3 = V / 64;
    h->V = V;
    STP_HIP(dalloc(&h->feat, V * NFEAT));
    STP_HIP(dalloc(&h->a, V * 16));
    STP_HIP(dalloc(&h->pa, V2 * 16));
    STP_HIP(dalloc(&h->b, V2 * 32));
    STP_HIP(dalloc(&h->pb, V3 * 32));
    STP_HIP(dalloc(&h->c, V3 * 64));
    STP_HIP(dalloc(&h->cat2, V2 * 96));
    STP_HIP(dalloc(&h->d, V2 * 32));
    STP_HIP(dalloc(&h->cat1, V * 48));
    STP_HIP(dalloc(&h->e, V * 16));
    STP_HIP(dalloc(&h->out, V * NOUT));
    STP_HIP(dalloc(&h->coef, V * NCOEF));
    STP_HIP(dalloc(&h->dsq, V));
    STP_HIP(dalloc(&h->nloc, V * NTAP));
    if (keep_train) {
        STP_HIP(dalloc(&h->dacc, V * NCOEF));
        STP_HIP(dalloc(&h->de, V * 16));
        STP_HIP(dalloc(&h->dcat1, V * 48));
        STP_HIP(dalloc(&h->dd, V2 * 32));
        STP_HIP(dalloc(&h->da, V * 16));
        STP_HIP(dalloc(&h->dcat2, V2 * 96));
        STP_HIP(dalloc(&h->dc, V3 * 64));
        STP_HIP(dalloc(&h->db, V2 * 32));
        STP_HIP(dalloc(&h->dpb, V3 * 32));
        STP_HIP(dalloc(&h->dpa, V2 * 16));
        STP_HIP(dalloc(&h->dO, V * NOUT));
        STP_HIP(dalloc(&h->dpre, V * 32));
        h->train_ws = true;
    }
    return STP_OK;
}

int check_system(const stp_csr* A, const int64_t* n) {
    if (!A || !n) return fail(STP_ERR_INVALID, "learned preconditioner: null argument");
    if ((int64_t)A->cols != n[0] * n[1] * n[2])
        return fail(STP_ERR_INVALID, "learned preconditioner: matrix columns do not match the grid");
    return STP_OK;
}

// features (and dsq) of system A on the current grid
int features(stp_lpc* h, const stp_csr* A, double lam, float* feat32, hipStream_t st) {
    const int64_t V = h->V;
    STP_HIP(hipMemsetAsync(h->nloc, 0, sizeof(double) * V * NTAP, st));
    if (V > 0)
        hipLaunchKernelGGL(couplings_kernel, dim3((unsigned)((V + 3) / 4)), dim3(256), 0, st, h->n0, h->n1, h->n2,
                           A->trp, A->tci, A->tval, h->nloc);
    STP_HIP(hipGetLastError());
    hipLaunchKernelGGL(features_kernel, dim3(gridn(V)), dim3(256), 0, st, h->n0, h->n1, h->n2, h->nloc, lam, h->feat,
                       h->dsq, feat32);
    STP_HIP(hipGetLastError());
    return STP_OK;
}

stp::ConvShape shp(int n0, int n1, int n2, int l) { return stp::ConvShape{n0, n1, n2, LAYERS[l].cin, LAYERS[l].cout}; }

int net_forward(stp_lpc* h, hipStream_t st) {
    const int a0 = h->n0, a1 = h->n1, a2 = h->n2;
    const int b0 = a0 / 2, b1 = a1 / 2, b2 = a2 / 2;
    const int c0 = a0 / 4, c1 = a1 / 4, c2 = a2 / 4;
    auto W = [&](int l) { return h->wbf + layer_off(l); };
    using namespace stp;
    STP_HIP(conv3d_fwd(shp(a0, a1, a2, E1), h->feat, W(E1), 1, CONV_OUT_BF16, h->a, 16, 0, st));
    STP_HIP(avgpool2_fwd(a0, a1, a2, 16, h->a, h->pa, st));
    STP_HIP(conv3d_fwd(shp(b0, b1, b2, E2), h->pa, W(E2), 1, CONV_OUT_BF16, h->b, 32, 0, st));
    STP_HIP(avgpool2_fwd(b0, b1, b2, 32, h->b, h->pb, st));
    STP_HIP(conv3d_fwd(shp(c0, c1, c2, E3), h->pb, W(E3), 1, CONV_OUT_BF16, h->c, 64, 0, st));
    STP_HIP(upcat_fwd(b0, b1, b2, 64, 32, h->c, h->b, h->cat2, st));
    STP_HIP(conv3d_fwd(shp(b0, b1, b2, D2), h->cat2, W(D2), 1, CONV_OUT_BF16, h->d, 32, 0, st));
    STP_HIP(upcat_fwd(a0, a1, a2, 32, 16, h->d, h->a, h->cat1, st));
    STP_HIP(conv3d_fwd(shp(a0, a1, a2, D1), h->cat1, W(D1), 1, CONV_OUT_BF16, h->e, 16, 0, st));
    STP_HIP(conv3d_fwd(shp(a0, a1, a2, LO), h->e, W(LO), 0, CONV_OUT_F32_SOA, h->out, 0, 0, st));
    hipLaunchKernelGGL(finalize_coef_kernel, dim3(gridn(h->V)), dim3(256), 0, st, a0, a1, a2, h->out, h->dsq,
                       h->coef);
    STP_HIP(hipGetLastError());
    return STP_OK;
}

// data gradient of layer l: din (FP32 channels-last, cin channels) from dpre (BF16, cout channels)
int conv_bwd_data(stp_lpc* h, int l, int n0, int n1, int n2, const bf16_t* dpre, float* din, hipStream_t st) {
    using namespace stp;
    STP_HIP(conv3d_flip_weights(LAYERS[l].cin, LAYERS[l].cout, h->w32 + layer_off(l), h->wflip, st));
    const ConvShape s{n0, n1, n2, LAYERS[l].cout, LAYERS[l].cin};
    STP_HIP(conv3d_fwd(s, dpre, h->wflip, 0, CONV_OUT_F32, din, LAYERS[l].cin, 0, st));
    return STP_OK;
}

// parameter gradient (into h->grad, overwritten) from dL/d(output) in h->dO
int net_backward(stp_lpc* h, hipStream_t st) {
    using namespace stp;
    const int a0 = h->n0, a1 = h->n1, a2 = h->n2;
    const int b0 = a0 / 2, b1 = a1 / 2, b2 = a2 / 2;
    const int c0 = a0 / 4, c1 = a1 / 4, c2 = a2 / 4;
    const int64_t V = h->V, V2 = V / 8, V3 = V / 64;
    auto G = [&](int l) { return h->grad + layer_off(l); };
    int rc;
    STP_HIP(hipMemsetAsync(h->grad, 0, sizeof(float) * total_params(), st));
    // o
    STP_HIP(conv3d_bwd_weight(shp(a0, a1, a2, LO), h->e, h->dO, G(LO), h->wdw, st));
    if ((rc = conv_bwd_data(h, LO, a0, a1, a2, h->dO, h->de, st)) != STP_OK) return rc;
    STP_HIP(relu_mask_bf16(V * 16, h->de, h->e, h->dpre, st));
    // d1
    STP_HIP(conv3d_bwd_weight(shp(a0, a1, a2, D1), h->cat1, h->dpre, G(D1), h->wdw, st));
    if ((rc = conv_bwd_data(h, D1, a0, a1, a2, h->dpre, h->dcat1, st)) != STP_OK) return rc;
    STP_HIP(upcat_bwd(a0, a1, a2, 32, 16, h->dcat1, h->dd, h->da, st));
    STP_HIP(relu_mask_bf16(V2 * 32, h->dd, h->d, h->dpre, st));
    // d2
    STP_HIP(conv3d_bwd_weight(shp(b0, b1, b2, D2), h->cat2, h->dpre, G(D2), h->wdw, st));
    if ((rc = conv_bwd_data(h, D2, b0, b1, b2, h->dpre, h->dcat2, st)) != STP_OK) return rc;
    STP_HIP(upcat_bwd(b0, b1, b2, 64, 32, h->dcat2, h->dc, h->db, st));
    STP_HIP(relu_mask_bf16(V3 * 64, h->dc, h->c, h->dpre, st));
    // e3
    STP_HIP(conv3d_bwd_weight(shp(c0, c1, c2, E3), h->pb, h->dpre, G(E3), h->wdw, st));
    if ((rc = conv_bwd_data(h, E3, c0, c1, c2, h->dpre, h->dpb, st)) != STP_OK) return rc;
    STP_HIP(avgpool2_bwd(b0, b1, b2, 32, h->dpb, h->db, st));
    STP_HIP(relu_mask_bf16(V2 * 32, h->db, h->b, h->dpre, st));
    // e2
    STP_HIP(conv3d_bwd_weight(shp(b0, b1, b2, E2), h->pa, h->dpre, G(E2), h->wdw, st));
    if ((rc = conv_bwd_data(h, E2, b0, b1, b2, h->dpre, h->dpa, st)) != STP_OK) return rc;
    STP_HIP(avgpool2_bwd(a0, a1, a2, 16, h->dpa, h->da, st));
    STP_HIP(relu_mask_bf16(V * 16, h->da, h->a, h->dpre, st));
    // e1
    STP_HIP(conv3d_bwd_weight(shp(a0, a1, a2, E1), h->feat, h->dpre, G(E1), h->wdw, st));
    return STP_OK;
}

}  // namespace

namespace stp {
int lpc_apply(void* handle, const float* s, float* z, const int* done, const float* s_prev, double* part_zs,
              double* part_fl, hipStream_t st, const int* perm) {
    stp_lpc* h = static_cast<stp_lpc*>(handle);
    if (!h || !h->prepared) return fail(STP_ERR_STATE, "learned preconditioner: not prepared for a system");
    if ((part_zs || part_fl) && !(part_zs && part_fl && s_prev))
        return fail(STP_ERR_INVALID, "learned preconditioner: incomplete dot-product outputs");
    const int64_t V = h->V;
    const dim3 grid((unsigned)((V + 255) / 256));
#define STP_APPLY_M(PERM, DOTS) hipLaunchKernelGGL((apply_m_kernel<PERM, DOTS>), grid, dim3(256), 0, st, h->n0, \
        h->n1, h->n2, h->coef, s, z, done, s_prev, part_zs, part_fl, perm)
    if (perm) {
        if (part_zs) STP_APPLY_M(true, true);
        else STP_APPLY_M(true, false);
    } else {
        if (part_zs) STP_APPLY_M(false, true);
        else STP_APPLY_M(false, false);
    }
#undef STP_APPLY_M
    STP_HIP(hipGetLastError());
    return STP_OK;
}
int64_t lpc_cols(const void* handle) {
    const stp_lpc* h = static_cast<const stp_lpc*>(handle);
    return h && h->prepared ? h->V : -1;
}
}  // namespace stp

extern "C" {

int stp_lpc_create(uint64_t seed, stp_lpc** out) {
    if (!out) return fail(STP_ERR_INVALID, "stp_lpc_create: null argument");
    *out = nullptr;
    stp_lpc* h = new (std::nothrow) stp_lpc();
    if (!h) return fail(STP_ERR_ALLOC, "stp_lpc_create: out of host memory");
    const int64_t P = total_params();
    int64_t maxl = 0;
    for (int l = 0; l < 6; ++l) maxl = std::max(maxl, layer_size(l));
    hipError_t e;
    if ((e = dalloc(&h->w32, P)) != hipSuccess || (e = dalloc(&h->wbf, P)) != hipSuccess ||
        (e = dalloc(&h->wflip, maxl)) != hipSuccess || (e = dalloc(&h->grad, P)) != hipSuccess ||
        (e = dalloc(&h->m1, P)) != hipSuccess || (e = dalloc(&h->m2, P)) != hipSuccess ||
        (e = dalloc(&h->wdw, stp::CONV_BWD_WORK_MAX)) != hipSuccess) {
        free_all(h);
        delete h;
        return stp::hip_fail(e, "stp_lpc_create");
    }
    // He-uniform initialisation over the real fan-in; the output layer starts
    // at zero (M = D^-1/2 (I + W0 L) D^-1/2), padded input channels stay zero.
    std::vector<float> w(P, 0.f);
    uint64_t s = seed;
    for (int l = 0; l < 6; ++l) {
        const int cin = LAYERS[l].cin, cout = LAYERS[l].cout;
        const int cin_real = l == E1 ? 27 : cin;
        const float bound = std::sqrt(6.0f / (27.0f * cin_real));
        for (int co = 0; co < cout; ++co)
            for (int t = 0; t < NTAP; ++t)
                for (int ci = 0; ci < cin; ++ci) {
                    const double u = (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
                    float v = (float)((2.0 * u - 1.0) * bound);
                    if (l == LO || ci >= cin_real) v = 0.f;
                    w[layer_off(l) + ((int64_t)co * NTAP + t) * cin + ci] = v;
                }
    }
    const int rc = stp_lpc_set_params(h, w.data(), P);
    if (rc != STP_OK) {
        free_all(h);
        delete h;
        return rc;
    }
    *out = h;
    return STP_OK;
}

int stp_lpc_destroy(stp_lpc* h) {
    if (!h) return STP_OK;
    free_all(h);
    delete h;
    return STP_OK;
}

int stp_lpc_num_params(int64_t* n) {
    if (!n) return fail(STP_ERR_INVALID, "stp_lpc_num_params: null argument");
    *n = total_params();
    return STP_OK;
}

int stp_lpc_get_params(const stp_lpc* h, float* host, int64_t n) {
    if (!h || !host || n != total_params()) return fail(STP_ERR_INVALID, "stp_lpc_get_params: bad argument");
    STP_HIP(hipMemcpy(host, h->w32, sizeof(float) * n, hipMemcpyDeviceToHost));
    return STP_OK;
}

int stp_lpc_set_params(stp_lpc* h, const float* host, int64_t n) {
    if (!h || !host || n != total_params()) return fail(STP_ERR_INVALID, "stp_lpc_set_params: bad argument");
    STP_HIP(hipMemcpy(h->w32, host, sizeof(float) * n, hipMemcpyHostToDevice));
    STP_HIP(stp::f32_to_bf16(n, h->w32, h->wbf, nullptr));
    STP_HIP(hipMemset(h->m1, 0, sizeof(float) * n));
    STP_HIP(hipMemset(h->m2, 0, sizeof(float) * n));
    STP_HIP(hipDeviceSynchronize());
    h->steps = 0;
    h->prepared = false;
    return STP_OK;
}

int stp_lpc_features(stp_lpc* h, const stp_csr* A, const int64_t* n, double lam, float* feat_dev, float* dsq_dev,
                     void* stream) {
    if (!h) return fail(STP_ERR_INVALID, "stp_lpc_features: null handle");
    int rc = check_system(A, n);
    if (rc != STP_OK) return rc;
    if ((rc = ensure_grid(h, n, false)) != STP_OK) return rc;
    hipStream_t st = S(stream);
    h->prepared = false;
    if ((rc = features(h, A, lam, feat_dev, st)) != STP_OK) return rc;
    if (dsq_dev) STP_HIP(hipMemcpyAsync(dsq_dev, h->dsq, sizeof(float) * h->V, hipMemcpyDeviceToDevice, st));
    return STP_OK;
}

static int prepare(stp_lpc* h, const stp_csr* A, const int64_t* n, double lam, bool train, void* stream) {
    if (!h) return fail(STP_ERR_INVALID, "stp_lpc_prepare: null handle");
    int rc = check_system(A, n);
    if (rc != STP_OK) return rc;
    if ((rc = ensure_grid(h, n, train)) != STP_OK) return rc;
    hipStream_t st = S(stream);
    h->prepared = false;
    if ((rc = features(h, A, lam, nullptr, st)) != STP_OK) return rc;
    if ((rc = net_forward(h, st)) != STP_OK) return rc;
    h->prepared = true;
    h->dacc_live = false;
    return STP_OK;
}

int stp_lpc_prepare(stp_lpc* h, const stp_csr* A, const int64_t* n, double lam, void* stream) {
    return prepare(h, A, n, lam, false, stream);
}

int stp_lpc_prepare_train(stp_lpc* h, const stp_csr* A, const int64_t* n, double lam, void* stream) {
    return prepare(h, A, n, lam, true, stream);
}

int stp_lpc_coefficients(const stp_lpc* h, float* coef_dev, void* stream) {
    if (!h || !coef_dev) return fail(STP_ERR_INVALID, "stp_lpc_coefficients: null argument");
    if (!h->prepared) return fail(STP_ERR_STATE, "stp_lpc_coefficients: not prepared for a system");
    STP_HIP(hipMemcpyAsync(coef_dev, h->coef, sizeof(float) * NCOEF * h->V, hipMemcpyDeviceToDevice, S(stream)));
    return STP_OK;
}

int stp_lpc_apply(const stp_lpc* h, const float* s_dev, float* z_dev, void* stream) {
    if (!h || !s_dev || !z_dev) return fail(STP_ERR_INVALID, "stp_lpc_apply: null argument");
    return stp::lpc_apply(const_cast<stp_lpc*>(h), s_dev, z_dev, nullptr, nullptr, nullptr, nullptr, S(stream));
}

int stp_lpc_backward_accumulate(stp_lpc* h, const float* r_dev, const float* g_dev, void* stream) {
    if (!h || !r_dev || !g_dev) return fail(STP_ERR_INVALID, "stp_lpc_backward_accumulate: null argument");
    if (!h->prepared) return fail(STP_ERR_STATE, "stp_lpc_backward_accumulate: not prepared for a system");
    hipStream_t st = S(stream);
    if (!h->train_ws)
        return fail(STP_ERR_STATE, "stp_lpc_backward_accumulate: system prepared without the training workspace "
                                   "(use stp_lpc_prepare_train)");
    if (!h->dacc_live) {
        STP_HIP(hipMemsetAsync(h->dacc, 0, sizeof(float) * NCOEF * h->V, st));
        h->dacc_live = true;
    }
    hipLaunchKernelGGL(apply_m_grad_kernel, dim3(gridn(h->V)), dim3(256), 0, st, h->n0, h->n1, h->n2, h->coef, r_dev,
                       g_dev, h->dacc);
    STP_HIP(hipGetLastError());
    return STP_OK;
}

int stp_lpc_backward(stp_lpc* h, double scale, float* grad_dev, void* stream) {
    if (!h) return fail(STP_ERR_INVALID, "stp_lpc_backward: null handle");
    if (!h->prepared || !h->train_ws || !h->dacc_live)
        return fail(STP_ERR_STATE, "stp_lpc_backward: no accumulated output gradient");
    hipStream_t st = S(stream);
    hipLaunchKernelGGL(out_grad_kernel, dim3(gridn(h->V * NOUT)), dim3(256), 0, st, h->V, h->dacc, h->coef, h->out,
                       (float)scale, h->dO);
    STP_HIP(hipGetLastError());
    const int rc = net_backward(h, st);
    if (rc != STP_OK) return rc;
    h->dacc_live = false;
    if (grad_dev)
        STP_HIP(hipMemcpyAsync(grad_dev, h->grad, sizeof(float) * total_params(), hipMemcpyDeviceToDevice, st));
    STP_HIP(hipStreamSynchronize(st));
    return STP_OK;
}

int stp_lpc_adam_step(stp_lpc* h, double lr, void* stream) {
    if (!h || !(lr > 0)) return fail(STP_ERR_INVALID, "stp_lpc_adam_step: bad argument");
    hipStream_t st = S(stream);
    const float b1 = 0.9f, b2 = 0.999f;
    h->steps += 1;
    const float c1 = 1.0f / (1.0f - std::pow(b1, (float)h->steps));
    const float c2 = 1.0f / (1.0f - std::pow(b2, (float)h->steps));
    STP_HIP(stp::adam_step(total_params(), h->w32, h->grad, h->m1, h->m2, (float)lr, b1, b2, 1e-8f, c1, c2, h->wbf, st));
    STP_HIP(hipStreamSynchronize(st));
    h->prepared = false;   // the coefficients no longer match the weights
    return STP_OK;
}

int stp_conv3d_bf16(int64_t n0, int64_t n1, int64_t n2, int32_t cin, int32_t cout, const uint16_t* in_dev,
                    const float* w_dev, int32_t relu, float* out_dev, void* stream) {
    if (!in_dev || !w_dev || !out_dev || n0 < 1 || n1 < 1 || n2 < 1 || n0 * n1 * n2 > (1 << 28) || (cin & 15) ||
        (cout & 15) || cin < 16 || cout < 16 || cout > 96 || cout == 80)
        return fail(STP_ERR_INVALID, "stp_conv3d_bf16: bad argument");
    hipStream_t st = S(stream);
    bf16_t* wb = nullptr;
    const int64_t nw = (int64_t)cin * cout * NTAP;
    STP_HIP(dalloc(&wb, nw));
    hipError_t e = stp::f32_to_bf16(nw, w_dev, wb, st);
    if (e == hipSuccess)
        e = stp::conv3d_fwd(stp::ConvShape{(int)n0, (int)n1, (int)n2, cin, cout}, in_dev, wb, relu, stp::CONV_OUT_F32,
                            out_dev, cout, 0, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(wb);
    if (e != hipSuccess) return stp::hip_fail(e, "stp_conv3d_bf16");
    return STP_OK;
}

int stp_conv3d_bf16_bwd(int64_t n0, int64_t n1, int64_t n2, int32_t cin, int32_t cout, const uint16_t* in_dev,
                        const float* w_dev, const uint16_t* dout_dev, float* din_dev, float* dw_dev, void* stream) {
    if (!in_dev || !w_dev || !dout_dev || n0 < 1 || n1 < 1 || n2 < 1 || n0 * n1 * n2 > (1 << 28) || (cin & 15) ||
        (cout & 15) || cin < 16 || cout < 16 || cin > 96 || cin == 80)
        return fail(STP_ERR_INVALID, "stp_conv3d_bf16_bwd: bad argument");
    hipStream_t st = S(stream);
    const stp::ConvShape s{(int)n0, (int)n1, (int)n2, cin, cout};
    hipError_t e = hipSuccess;
    if (dw_dev) {
        float* work = nullptr;
        e = dalloc(&work, std::max<size_t>(stp::conv3d_bwd_weight_work(s), 1));
        if (e == hipSuccess) e = stp::conv3d_bwd_weight(s, in_dev, dout_dev, dw_dev, work, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (work) (void)hipFree(work);
    }
    if (e == hipSuccess && din_dev) {
        bf16_t* wb = nullptr;
        e = dalloc(&wb, (size_t)cin * cout * NTAP);
        if (e == hipSuccess) e = stp::conv3d_flip_weights(cin, cout, w_dev, wb, st);
        if (e == hipSuccess)
            e = stp::conv3d_fwd(stp::ConvShape{(int)n0, (int)n1, (int)n2, cout, cin}, dout_dev, wb, 0,
                                stp::CONV_OUT_F32, din_dev, cin, 0, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (wb) (void)hipFree(wb);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return stp::hip_fail(e, "stp_conv3d_bf16_bwd");
    return STP_OK;
}

}  // extern "C"
