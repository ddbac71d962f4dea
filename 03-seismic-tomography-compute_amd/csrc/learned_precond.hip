// Learned preconditioner for CGLS on the ray-tomography normal equations
// N = A^T A + lam I (declarations: include/seismic_hip.h).
//
// The network reads the system, not the residual.  For each cell it gets the
// Jacobi-scaled couplings of N to its 26 neighbours,
//   J[i][i+d] = N[i][i+d] / sqrt(N[i][i] N[i+d][i+d]),
// plus a constant channel (which marks the model boundary through the zero
// padding), and it outputs the 63 coefficients of row i of an upper-triangular
// factor G (taps d in [-2,2]^3 lexicographically >= 0, the centre first; the
// diagonal is exp(output) > 0).  The preconditioner is
//   M = D^-1/2 G G^T D^-1/2,   D = diag(N),
// symmetric positive definite for any weights, so the CG theory holds; the
// network runs once per system and each iteration costs two 63-tap stencils.
//
// Training minimises Kaporin's condition number of G^T J G
// (J = D^-1/2 N D^-1/2):
//   K = (tr(G^T J G) / n) / det(G^T J G)^(1/n),
//   log K = log(tr(G^T J G) / n) - (2/n) sum_i log G_ii - (1/n) log det J,
// with the trace estimated from Rademacher probe vectors y, E[y y^T] = I:
//   tr(G^T J G) ~ (1/P) sum_y (G y)^T J (G y).
// K >= 1 with equality iff G G^T is a multiple of J^-1, it bounds the CG
// iteration count (Kaporin 1994), and it is invariant to the scale of G.  Its
// minimiser over a fixed pattern is the factorised sparse approximate inverse
// (FSAI) of J; the network predicts such a factor from local couplings.
//
// Network: 3 levels, 16/32/64 channels, 3x3x3 convolutions, ReLU, average
// pooling down, nearest upsampling and skip concatenation up (conv3d.hip),
// BF16 weights and activations, FP32 accumulation and FP32 master weights
// (Adam).  Layers, [cout][27][cin]:
//   e1 32->16, e2 16->32 (after pool), e3 32->64 (after pool),
//   d2 (64+32)->32, d1 (32+16)->16, o 16->64 (63 used, no ReLU).
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
constexpr int RG = 2;       // stencil radius of the factor G
constexpr int NG = 63;      // taps of G: offsets in [-2,2]^3 lexicographically >= 0
constexpr int NOUT = 64;    // network output channels (NG used)
constexpr int PARTS = 1024; // FP64 block partial sums per reduction

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
// One wavefront per row of A (a ray).  Columns of a row are sorted, so the
// neighbours of cell i in the row follow it within n1*n2 + n2 + 1 of its index.
__global__ void __launch_bounds__(256) couplings_kernel(int rows, const int* __restrict__ rp,
                                                        const int* __restrict__ ci, const float* __restrict__ val,
                                                        int n0, int n1, int n2, double* nloc) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= rows) return;
    const int lane = threadIdx.x & 63;
    const int beg = rp[row], end = rp[row + 1];
    const int64_t V = (int64_t)n0 * n1 * n2;
    const int reach = n1 * n2 + n2 + 1;
    for (int k = beg + lane; k < end; k += 64) {
        const int i = ci[k];
        const double a = val[k];
        const int zi = i % n2, yi = (i / n2) % n1, xi = i / (n1 * n2);
        unsafeAtomicAdd(nloc + 13 * V + i, a * a);
        for (int k2 = k + 1; k2 < end; ++k2) {
            const int j = ci[k2];
            if (j - i > reach) break;
            const int dz = j % n2 - zi, dy = (j / n2) % n1 - yi, dx = j / (n1 * n2) - xi;
            if (dx < -1 || dx > 1 || dy < -1 || dy > 1 || dz < -1 || dz > 1) continue;
            const int s = (dx + 1) * 9 + (dy + 1) * 3 + (dz + 1);
            const double pr = a * (double)val[k2];
            unsafeAtomicAdd(nloc + s * V + i, pr);
            unsafeAtomicAdd(nloc + (26 - s) * V + j, pr);
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

// ------------------------------------------------------------------ stencils of G
// Tap k of G couples cell i to cell i + d_k, d_k = the (62 + k)-th offset of
// [-2,2]^3 in lexicographic (x, y, z) order; k = 0 is the diagonal.
__device__ __forceinline__ void g_offset(int k, int& dx, int& dy, int& dz) {
    const int l = 62 + k;
    dx = l / 25 - RG;
    dy = (l / 5) % 5 - RG;
    dz = l % 5 - RG;
}

// u = sy .* (G (sx .* y)):  u_i = sum_k g[k][i] x_{i + d_k}   (sx, sy may be null = 1)
__global__ void __launch_bounds__(256) apply_g_kernel(int n0, int n1, int n2, const float* __restrict__ g,
                                                      const float* __restrict__ y, const float* __restrict__ sx,
                                                      const float* __restrict__ sy, float* __restrict__ u,
                                                      const int* done) {
    if (done && *done) return;
    const int64_t V = (int64_t)n0 * n1 * n2;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < V; i += (int64_t)gridDim.x * blockDim.x) {
        const int z = (int)(i % n2), yv = (int)((i / n2) % n1), x = (int)(i / ((int64_t)n1 * n2));
        float acc = 0.f;
#pragma unroll
        for (int k = 0; k < NG; ++k) {
            int dx, dy, dz;
            g_offset(k, dx, dy, dz);
            const int xx = x + dx, yy = yv + dy, zz = z + dz;
            if ((unsigned)xx < (unsigned)n0 && (unsigned)yy < (unsigned)n1 && (unsigned)zz < (unsigned)n2) {
                const int64_t j = ((int64_t)xx * n1 + yy) * n2 + zz;
                acc += g[k * V + i] * y[j] * (sx ? sx[j] : 1.f);
            }
        }
        u[i] = sy ? sy[i] * acc : acc;
    }
}

// w = sy .* (G^T (sx .* v)):  w_j = sum_k g[k][j - d_k] x_{j - d_k}
__global__ void __launch_bounds__(256) apply_gt_kernel(int n0, int n1, int n2, const float* __restrict__ g,
                                                       const float* __restrict__ v, const float* __restrict__ sx,
                                                       const float* __restrict__ sy, float* __restrict__ w,
                                                       const int* done) {
    if (done && *done) return;
    const int64_t V = (int64_t)n0 * n1 * n2;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < V; j += (int64_t)gridDim.x * blockDim.x) {
        const int z = (int)(j % n2), yv = (int)((j / n2) % n1), x = (int)(j / ((int64_t)n1 * n2));
        float acc = 0.f;
#pragma unroll
        for (int k = 0; k < NG; ++k) {
            int dx, dy, dz;
            g_offset(k, dx, dy, dz);
            const int xx = x - dx, yy = yv - dy, zz = z - dz;
            if ((unsigned)xx < (unsigned)n0 && (unsigned)yy < (unsigned)n1 && (unsigned)zz < (unsigned)n2) {
                const int64_t i = ((int64_t)xx * n1 + yy) * n2 + zz;
                acc += g[k * V + i] * v[i] * (sx ? sx[i] : 1.f);
            }
        }
        w[j] = sy ? sy[j] * acc : acc;
    }
}

// dg[k][i] += a_i b_{i + d_k}
__global__ void __launch_bounds__(256) coef_grad_kernel(int n0, int n1, int n2, const float* __restrict__ a,
                                                        const float* __restrict__ b, float* __restrict__ dg) {
    const int64_t V = (int64_t)n0 * n1 * n2;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < V; i += (int64_t)gridDim.x * blockDim.x) {
        const int z = (int)(i % n2), yv = (int)((i / n2) % n1), x = (int)(i / ((int64_t)n1 * n2));
        const float ai = a[i];
#pragma unroll
        for (int k = 0; k < NG; ++k) {
            int dx, dy, dz;
            g_offset(k, dx, dy, dz);
            const int xx = x + dx, yy = yv + dy, zz = z + dz;
            if ((unsigned)xx < (unsigned)n0 && (unsigned)yy < (unsigned)n1 && (unsigned)zz < (unsigned)n2) {
                const int64_t j = ((int64_t)xx * n1 + yy) * n2 + zz;
                dg[k * V + i] += ai * b[j];
            }
        }
    }
}

// network output (SoA, NOUT rows) -> coefficients: diagonal exp(o), clamped
// to exp(+-30); FP64 partial sums of the clamped exponent (= log G_ii)
__global__ void finalize_g_kernel(int64_t V, float* g, double* part) {
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < V; i += (int64_t)gridDim.x * blockDim.x) {
        const float o = fminf(fmaxf(g[i], -30.f), 30.f);
        g[i] = expf(o);
        acc += o;
    }
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0 && part) part[blockIdx.x] = red[0];
}

// out = x .* s ; out2 = s .* (t + lam * x .* s)   (two helpers of J)
__global__ void scale_kernel(int64_t n, const float* x, const float* s, float* out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = x[i] * s[i];
}
__global__ void jfinish_kernel(int64_t n, const float* t, const float* xs, const float* s, float lam, float* out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = s[i] * (t[i] + lam * xs[i]);
}

// FP64 partial sums of a . b (one per block of 256)
__global__ void dot_part_kernel(int64_t n, const float* a, const float* b, double* part) {
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        acc += (double)a[i] * (double)b[i];
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// gradient of log K with respect to the network output, channels-last BF16:
//   channel k >= 1: coef * dg[k][p];  channel 0: coef * dg[0][p] * G_pp - 2/V;
//   channels >= NG: 0
__global__ void out_grad_kernel(int64_t V, const float* dg, const float* g, float coef, float diag, bf16_t* d) {
    const int64_t n = V * NOUT;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(k % NOUT);
        const int64_t p = k / NOUT;
        float v = 0.f;
        if (c == 0)
            v = coef * dg[p] * g[p] - diag;
        else if (c < NG)
            v = coef * dg[c * V + p];
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
    int64_t steps = 0;
    // grid-sized workspace
    int n0 = 0, n1 = 0, n2 = 0;
    int64_t V = 0;
    bool train_ws = false;
    bf16_t *feat = nullptr, *a = nullptr, *pa = nullptr, *b = nullptr, *pb = nullptr, *c = nullptr, *cat2 = nullptr,
           *d = nullptr, *cat1 = nullptr, *e = nullptr;
    float *g = nullptr, *dsq = nullptr, *tmp = nullptr;
    double* nloc = nullptr;
    // training workspace
    float *dg = nullptr, *de = nullptr, *dcat1 = nullptr, *dd = nullptr, *da = nullptr, *dcat2 = nullptr,
          *dc = nullptr, *db = nullptr, *dpb = nullptr, *dpa = nullptr;
    bf16_t *dO = nullptr, *dpre = nullptr;
    float *u = nullptr, *w1 = nullptr, *t = nullptr, *v = nullptr, *q = nullptr;
    double* part = nullptr;   // 2 x PARTS FP64 block partial sums
    int64_t qcap = 0;
    // state of the prepared preconditioner
    bool prepared = false;
};

namespace {

void free_grid(stp_lpc* h) {
    void* ps[] = {h->feat, h->a,  h->pa,    h->b,  h->pb, h->c,  h->cat2, h->d,   h->cat1, h->e,  h->g,
                  h->dsq,  h->tmp, h->nloc, h->dg, h->de, h->dcat1, h->dd, h->da, h->dcat2, h->dc, h->db,
                  h->dpb,  h->dpa, h->dO,   h->dpre, h->u, h->w1, h->t,  h->v,  h->part};
    for (void* p : ps)
        if (p) hipFree(p);
    h->feat = h->a = h->pa = h->b = h->pb = h->c = h->cat2 = h->d = h->cat1 = h->e = nullptr;
    h->g = h->dsq = h->tmp = nullptr;
    h->nloc = nullptr;
    h->dg = h->de = h->dcat1 = h->dd = h->da = h->dcat2 = h->dc = h->db = h->dpb = h->dpa = nullptr;
    h->dO = h->dpre = nullptr;
    h->u = h->w1 = h->t = h->v = nullptr;
    h->part = nullptr;
    h->V = 0;
    h->n0 = h->n1 = h->n2 = 0;
    h->train_ws = false;
    h->prepared = false;
}

void free_all(stp_lpc* h) {
    free_grid(h);
    void* ps[] = {h->w32, h->wbf, h->wflip, h->grad, h->m1, h->m2, h->q};
    for (void* p : ps)
        if (p) hipFree(p);
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
    STP_HIP(dalloc(&h->g, V * NOUT));
    STP_HIP(dalloc(&h->part, 2 * PARTS));
    STP_HIP(dalloc(&h->dsq, V));
    STP_HIP(dalloc(&h->tmp, V));
    STP_HIP(dalloc(&h->nloc, V * NTAP));
    if (keep_train) {
        STP_HIP(dalloc(&h->dg, V * NG));
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
        STP_HIP(dalloc(&h->u, V));
        STP_HIP(dalloc(&h->w1, V));
        STP_HIP(dalloc(&h->t, V));
        STP_HIP(dalloc(&h->v, V));
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
    if (A->rows > 0)
        hipLaunchKernelGGL(couplings_kernel, dim3((A->rows + 3) / 4), dim3(256), 0, st, A->rows, A->rp, A->ci, A->val,
                           h->n0, h->n1, h->n2, h->nloc);
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
    STP_HIP(conv3d_fwd(shp(a0, a1, a2, LO), h->e, W(LO), 0, CONV_OUT_F32_SOA, h->g, 0, 0, st));
    // diagonal = exp(output 0); partial sums of log G_ii into part[PARTS ..]
    hipLaunchKernelGGL(finalize_g_kernel, dim3(PARTS), dim3(256), 0, st, h->V, h->g, h->part + PARTS);
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
    STP_HIP(conv3d_bwd_weight(shp(a0, a1, a2, LO), h->e, h->dO, G(LO), st));
    if ((rc = conv_bwd_data(h, LO, a0, a1, a2, h->dO, h->de, st)) != STP_OK) return rc;
    STP_HIP(relu_mask_bf16(V * 16, h->de, h->e, h->dpre, st));
    // d1
    STP_HIP(conv3d_bwd_weight(shp(a0, a1, a2, D1), h->cat1, h->dpre, G(D1), st));
    if ((rc = conv_bwd_data(h, D1, a0, a1, a2, h->dpre, h->dcat1, st)) != STP_OK) return rc;
    STP_HIP(upcat_bwd(a0, a1, a2, 32, 16, h->dcat1, h->dd, h->da, st));
    STP_HIP(relu_mask_bf16(V2 * 32, h->dd, h->d, h->dpre, st));
    // d2
    STP_HIP(conv3d_bwd_weight(shp(b0, b1, b2, D2), h->cat2, h->dpre, G(D2), st));
    if ((rc = conv_bwd_data(h, D2, b0, b1, b2, h->dpre, h->dcat2, st)) != STP_OK) return rc;
    STP_HIP(upcat_bwd(b0, b1, b2, 64, 32, h->dcat2, h->dc, h->db, st));
    STP_HIP(relu_mask_bf16(V3 * 64, h->dc, h->c, h->dpre, st));
    // e3
    STP_HIP(conv3d_bwd_weight(shp(c0, c1, c2, E3), h->pb, h->dpre, G(E3), st));
    if ((rc = conv_bwd_data(h, E3, c0, c1, c2, h->dpre, h->dpb, st)) != STP_OK) return rc;
    STP_HIP(avgpool2_bwd(b0, b1, b2, 32, h->dpb, h->db, st));
    STP_HIP(relu_mask_bf16(V2 * 32, h->db, h->b, h->dpre, st));
    // e2
    STP_HIP(conv3d_bwd_weight(shp(b0, b1, b2, E2), h->pa, h->dpre, G(E2), st));
    if ((rc = conv_bwd_data(h, E2, b0, b1, b2, h->dpre, h->dpa, st)) != STP_OK) return rc;
    STP_HIP(avgpool2_bwd(a0, a1, a2, 16, h->dpa, h->da, st));
    STP_HIP(relu_mask_bf16(V * 16, h->da, h->a, h->dpre, st));
    // e1
    STP_HIP(conv3d_bwd_weight(shp(a0, a1, a2, E1), h->feat, h->dpre, G(E1), st));
    return STP_OK;
}

// J x = dsq .* (A^T A (dsq .* x) + lam dsq .* x)
int apply_j(stp_lpc* h, const stp_csr* A, double lam, const float* x, float* out, hipStream_t st) {
    const int64_t V = h->V;
    hipLaunchKernelGGL(scale_kernel, dim3(gridn(V)), dim3(256), 0, st, V, x, h->dsq, h->w1);
    STP_HIP(hipGetLastError());
    STP_HIP(stp::spmv_f32(A->W, A->rows, A->rp, A->ci, A->val, h->w1, h->q, nullptr, nullptr, st));
    STP_HIP(stp::spmv_f32(A->WT, A->cols, A->trp, A->tci, A->tval, h->q, h->t, nullptr, nullptr, st));
    hipLaunchKernelGGL(jfinish_kernel, dim3(gridn(V)), dim3(256), 0, st, V, h->t, h->w1, h->dsq, (float)lam, out);
    STP_HIP(hipGetLastError());
    return STP_OK;
}

// features + forward + loss + backward (gradient in h->grad).  Loss:
//   log K~ = log(q / (P V)) - (2/V) sum_i log G_ii,   q = sum_y (G y)^T J (G y)
// (the constant -(1/V) log det J is left out); dlogK/dg[k][i] =
// (2/q) sum_y (J G y)_i y_{i+d_k}, and through G_ii = exp(o_i) the diagonal
// output gets (2/q) sum_y (J G y)_i y_i G_ii - 2/V.
int loss_grad(stp_lpc* h, const stp_csr* A, const int64_t* n, double lam, const float* probes, int nprobe,
              double* loss, hipStream_t st) {
    int rc = check_system(A, n);
    if (rc != STP_OK) return rc;
    if (!probes || nprobe < 1) return fail(STP_ERR_INVALID, "learned preconditioner: no probe vectors");
    if ((rc = ensure_grid(h, n, true)) != STP_OK) return rc;
    if (h->qcap < A->rows) {
        if (h->q) hipFree(h->q);
        h->q = nullptr;
        h->qcap = 0;
        STP_HIP(dalloc(&h->q, A->rows));
        h->qcap = A->rows;
    }
    h->prepared = false;
    if ((rc = features(h, A, lam, nullptr, st)) != STP_OK) return rc;
    if ((rc = net_forward(h, st)) != STP_OK) return rc;
    const int64_t V = h->V;
    const int n0 = h->n0, n1 = h->n1, n2 = h->n2;
    STP_HIP(hipMemsetAsync(h->dg, 0, sizeof(float) * V * NG, st));
    std::vector<double> parts((size_t)(nprobe + 1) * PARTS);
    for (int k = 0; k < nprobe; ++k) {
        const float* y = probes + (int64_t)k * V;
        hipLaunchKernelGGL(apply_g_kernel, dim3(gridn(V)), dim3(256), 0, st, n0, n1, n2, h->g, y, nullptr, nullptr,
                           h->u, nullptr);
        STP_HIP(hipGetLastError());
        if ((rc = apply_j(h, A, lam, h->u, h->v, st)) != STP_OK) return rc;
        hipLaunchKernelGGL(dot_part_kernel, dim3(PARTS), dim3(256), 0, st, V, h->u, h->v, h->part);
        hipLaunchKernelGGL(coef_grad_kernel, dim3(gridn(V)), dim3(256), 0, st, n0, n1, n2, h->v, y, h->dg);
        STP_HIP(hipGetLastError());
        STP_HIP(hipMemcpyAsync(parts.data() + (size_t)k * PARTS, h->part, sizeof(double) * PARTS,
                               hipMemcpyDeviceToHost, st));
    }
    STP_HIP(hipMemcpyAsync(parts.data() + (size_t)nprobe * PARTS, h->part + PARTS, sizeof(double) * PARTS,
                           hipMemcpyDeviceToHost, st));
    STP_HIP(hipStreamSynchronize(st));
    double q = 0.0, logdiag = 0.0;
    for (int k = 0; k < nprobe; ++k)
        for (int j = 0; j < PARTS; ++j) q += parts[(size_t)k * PARTS + j];
    for (int j = 0; j < PARTS; ++j) logdiag += parts[(size_t)nprobe * PARTS + j];
    if (!(q > 0.0) || !std::isfinite(q))
        return fail(STP_ERR_INVALID, "learned preconditioner: non-positive or non-finite trace estimate");
    if (loss) *loss = std::log(q / ((double)nprobe * (double)V)) - 2.0 * logdiag / (double)V;
    hipLaunchKernelGGL(out_grad_kernel, dim3(gridn(V * NOUT)), dim3(256), 0, st, V, h->dg, h->g, (float)(2.0 / q),
                       (float)(2.0 / (double)V), h->dO);
    STP_HIP(hipGetLastError());
    if ((rc = net_backward(h, st)) != STP_OK) return rc;
    STP_HIP(hipStreamSynchronize(st));
    return STP_OK;
}

}  // namespace

namespace stp {
int lpc_apply(void* handle, const float* s, float* z, const int* done, hipStream_t st) {
    stp_lpc* h = static_cast<stp_lpc*>(handle);
    if (!h || !h->prepared) return fail(STP_ERR_STATE, "learned preconditioner: not prepared for a system");
    const int64_t V = h->V;
    // tmp = G^T (dsq .* s) ; z = dsq .* (G tmp)
    hipLaunchKernelGGL(apply_gt_kernel, dim3(gridn(V)), dim3(256), 0, st, h->n0, h->n1, h->n2, h->g, s, h->dsq,
                       nullptr, h->tmp, done);
    hipLaunchKernelGGL(apply_g_kernel, dim3(gridn(V)), dim3(256), 0, st, h->n0, h->n1, h->n2, h->g, h->tmp, nullptr,
                       h->dsq, z, done);
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
        (e = dalloc(&h->m1, P)) != hipSuccess || (e = dalloc(&h->m2, P)) != hipSuccess) {
        free_all(h);
        delete h;
        return stp::hip_fail(e, "stp_lpc_create");
    }
    // He-uniform initialisation over the real fan-in; the output layer starts
    // at zero (G = I, i.e. Jacobi), padded input/output channels stay zero.
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

int stp_lpc_prepare(stp_lpc* h, const stp_csr* A, const int64_t* n, double lam, void* stream) {
    if (!h) return fail(STP_ERR_INVALID, "stp_lpc_prepare: null handle");
    int rc = check_system(A, n);
    if (rc != STP_OK) return rc;
    if ((rc = ensure_grid(h, n, false)) != STP_OK) return rc;
    hipStream_t st = S(stream);
    h->prepared = false;
    if ((rc = features(h, A, lam, nullptr, st)) != STP_OK) return rc;
    if ((rc = net_forward(h, st)) != STP_OK) return rc;
    h->prepared = true;
    return STP_OK;
}

int stp_lpc_coefficients(const stp_lpc* h, float* g_dev, void* stream) {
    if (!h || !g_dev) return fail(STP_ERR_INVALID, "stp_lpc_coefficients: null argument");
    if (!h->prepared) return fail(STP_ERR_STATE, "stp_lpc_coefficients: not prepared for a system");
    STP_HIP(hipMemcpyAsync(g_dev, h->g, sizeof(float) * NG * h->V, hipMemcpyDeviceToDevice, S(stream)));
    return STP_OK;
}

int stp_lpc_apply(const stp_lpc* h, const float* s_dev, float* z_dev, void* stream) {
    if (!h || !s_dev || !z_dev) return fail(STP_ERR_INVALID, "stp_lpc_apply: null argument");
    return stp::lpc_apply(const_cast<stp_lpc*>(h), s_dev, z_dev, nullptr, S(stream));
}

int stp_lpc_loss_grad(stp_lpc* h, const stp_csr* A, const int64_t* n, double lam, const float* probes_dev,
                      int32_t nprobe, double* loss, float* grad_dev, void* stream) {
    if (!h) return fail(STP_ERR_INVALID, "stp_lpc_loss_grad: null handle");
    hipStream_t st = S(stream);
    int rc = loss_grad(h, A, n, lam, probes_dev, nprobe, loss, st);
    if (rc != STP_OK) return rc;
    if (grad_dev) {
        STP_HIP(hipMemcpyAsync(grad_dev, h->grad, sizeof(float) * total_params(), hipMemcpyDeviceToDevice, st));
        STP_HIP(hipStreamSynchronize(st));
    }
    return STP_OK;
}

int stp_lpc_train_step(stp_lpc* h, const stp_csr* A, const int64_t* n, double lam, const float* probes_dev,
                       int32_t nprobe, double lr, double* loss, void* stream) {
    if (!h || !(lr > 0)) return fail(STP_ERR_INVALID, "stp_lpc_train_step: bad argument");
    hipStream_t st = S(stream);
    int rc = loss_grad(h, A, n, lam, probes_dev, nprobe, loss, st);
    if (rc != STP_OK) return rc;
    const float b1 = 0.9f, b2 = 0.999f;
    h->steps += 1;
    const float c1 = 1.0f / (1.0f - std::pow(b1, (float)h->steps));
    const float c2 = 1.0f / (1.0f - std::pow(b2, (float)h->steps));
    STP_HIP(stp::adam_step(total_params(), h->w32, h->grad, h->m1, h->m2, (float)lr, b1, b2, 1e-8f, c1, c2, h->wbf, st));
    STP_HIP(hipStreamSynchronize(st));
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
    hipFree(wb);
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
        e = hipMemsetAsync(dw_dev, 0, sizeof(float) * cin * cout * NTAP, st);
        if (e == hipSuccess) e = stp::conv3d_bwd_weight(s, in_dev, dout_dev, dw_dev, st);
    }
    if (e == hipSuccess && din_dev) {
        bf16_t* wb = nullptr;
        e = dalloc(&wb, (size_t)cin * cout * NTAP);
        if (e == hipSuccess) e = stp::conv3d_flip_weights(cin, cout, w_dev, wb, st);
        if (e == hipSuccess)
            e = stp::conv3d_fwd(stp::ConvShape{(int)n0, (int)n1, (int)n2, cout, cin}, dout_dev, wb, 0,
                                stp::CONV_OUT_F32, din_dev, cin, 0, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (wb) hipFree(wb);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return stp::hip_fail(e, "stp_conv3d_bf16_bwd");
    return STP_OK;
}

}  // extern "C"
