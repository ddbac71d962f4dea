// 3x3x3 convolutions on MFMA (v_mfma_f32_16x16x32_bf16) for the learned
// preconditioner network, plus the pooling / upsampling / optimiser kernels
// around them (declarations: conv3d.h).
//
// Forward and data-gradient convolutions are one implicit GEMM:
//   out[p][co] = sum_{t, ci} in[p + d_t][ci] * w[co][t][ci]
// with M = voxels, N = cout, K = 27*cin.  A wavefront owns 16 consecutive
// voxels and every output channel (cout/16 accumulators); each K step of 32
// is one MFMA per 16 output channels.  Operand lane layout of the 16x16x32
// BF16 MFMA: lane l supplies A[l%16][8*(l/16) .. +7] and B[8*(l/16) .. +7][l%16]
// and receives D[4*(l/16) + i][l%16], i = 0..3.  Because cin is a multiple of
// 16, every group of 8 consecutive K indices is 8 consecutive channels of one
// tap: one 16-byte load per operand per lane.
//
// The weight gradient is the GEMM dw[co][(t,ci)] = sum_p dout[p][co] *
// in[p + d_t][ci] with K = voxels: a wavefront owns one 16x16 block of dw and
// a chunk of voxels and writes its partial block to a workspace; a second
// kernel sums the chunks in chunk order, so training is reproducible.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "conv3d.h"

namespace stp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(bf16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ bf16_t f2bf(float f) {
    __bf16 b = (__bf16)f;
    return __builtin_bit_cast(bf16_t, b);
}

union Pack8 {
    uint4 u;
    bf16x8 v;
    bf16_t h[8];
};

__device__ __forceinline__ void tap_offset(int t, int& dx, int& dy, int& dz) {
    dx = t / 9 - 1;
    dy = (t / 3) % 3 - 1;
    dz = t % 3 - 1;
}

template <int NB, int OUT>
__global__ void __launch_bounds__(256) conv_fwd_kernel(int n0, int n1, int n2, int cin, const bf16_t* __restrict__ in,
                                                       const bf16_t* __restrict__ w, int relu, void* out, int out_ld,
                                                       int out_off) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nvox = n0 * n1 * n2;
    const int tile = (blockIdx.x * 4 + wave) * 16;
    if (tile >= nvox) return;   // whole wavefront; no block-level synchronisation below
    const int m = lane & 15, kg = lane >> 4;
    const int p = tile + m;
    const bool pv = p < nvox;
    int x = 0, y = 0, z = 0;
    if (pv) {
        z = p % n2;
        const int q = p / n2;
        y = q % n1;
        x = q / n1;
    }
    const int K = 27 * cin;
    const int ngroups = K >> 3;
    const int nchunks = (ngroups + 3) >> 2;
    f32x4 acc[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[b] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int c = 0; c < nchunks; ++c) {
        const int g = c * 4 + kg;
        Pack8 a, bw[NB];
        a.u = make_uint4(0, 0, 0, 0);
#pragma unroll
        for (int b = 0; b < NB; ++b) bw[b].u = make_uint4(0, 0, 0, 0);
        if (g < ngroups) {
            const int k0 = g << 3;
            const int t = k0 / cin;
            const int ci0 = k0 - t * cin;
            int dx, dy, dz;
            tap_offset(t, dx, dy, dz);
            const int xx = x + dx, yy = y + dy, zz = z + dz;
            if (pv && (unsigned)xx < (unsigned)n0 && (unsigned)yy < (unsigned)n1 && (unsigned)zz < (unsigned)n2)
                a.u = *reinterpret_cast<const uint4*>(in + (size_t)((xx * n1 + yy) * n2 + zz) * cin + ci0);
#pragma unroll
            for (int b = 0; b < NB; ++b)
                bw[b].u = *reinterpret_cast<const uint4*>(w + (size_t)(b * 16 + m) * K + k0);
        }
#pragma unroll
        for (int b = 0; b < NB; ++b) acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v, bw[b].v, acc[b], 0, 0, 0);
    }
    // lane holds D[4*kg + i][m]: voxel tile + 4*kg + i, channel b*16 + m
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int co = b * 16 + m;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int pp = tile + 4 * kg + i;
            if (pp >= nvox) continue;
            float v = acc[b][i];
            if (relu) v = fmaxf(v, 0.f);
            if (OUT == CONV_OUT_BF16)
                static_cast<bf16_t*>(out)[(size_t)pp * out_ld + out_off + co] = f2bf(v);
            else if (OUT == CONV_OUT_F32)
                static_cast<float*>(out)[(size_t)pp * out_ld + out_off + co] = v;
            else
                static_cast<float*>(out)[(size_t)co * nvox + pp] = v;
        }
    }
}

// One wavefront per (16 x 16 block of dw, voxel chunk).  blockIdx.x: 4 dw
// blocks (one per wavefront), blockIdx.y: voxel chunk of `vc` voxels; each
// writes its chunk's partial block to work[chunk][co][t][ci].
__global__ void __launch_bounds__(256) conv_bwd_weight_kernel(int n0, int n1, int n2, int cin, int cout,
                                                              const bf16_t* __restrict__ in,
                                                              const bf16_t* __restrict__ dout,
                                                              float* __restrict__ work, int vc) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int ncolb = (27 * cin) >> 4;
    const int nblk = (cout >> 4) * ncolb;
    const int blk = blockIdx.x * 4 + wave;
    if (blk >= nblk) return;
    const int mb = blk / ncolb, cb = blk - mb * ncolb;
    const int m = lane & 15, kg = lane >> 4;
    const int co = mb * 16 + m;            // A row of this lane
    const int col = cb * 16 + m;           // B column of this lane
    const int t = col / cin, ci = col - t * cin;
    int dx, dy, dz;
    tap_offset(t, dx, dy, dz);
    const int nvox = n0 * n1 * n2;
    const int v0 = blockIdx.y * vc;
    const int v1 = min(nvox, v0 + vc);
    f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int base = v0; base < v1; base += 32) {
        Pack8 a, b;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int p = base + kg * 8 + j;
            bf16_t av = 0, bv = 0;
            if (p < v1) {
                av = dout[(size_t)p * cout + co];
                const int z = p % n2;
                const int q = p / n2;
                const int y = q % n1;
                const int x = q / n1;
                const int xx = x + dx, yy = y + dy, zz = z + dz;
                if ((unsigned)xx < (unsigned)n0 && (unsigned)yy < (unsigned)n1 && (unsigned)zz < (unsigned)n2)
                    bv = in[(size_t)((xx * n1 + yy) * n2 + zz) * cin + ci];
            }
            a.h[j] = av;
            b.h[j] = bv;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v, b.v, acc, 0, 0, 0);
    }
    // D[4*kg + i][m]: row co' = mb*16 + 4*kg + i, column cb*16 + m
    float* part = work + (size_t)blockIdx.y * cout * 27 * cin;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int cr = mb * 16 + 4 * kg + i;
        part[(size_t)cr * 27 * cin + cb * 16 + m] = acc[i];
    }
}

// dw[e] = sum over chunks c (in order) of work[c][e]
__global__ void sum_chunks_kernel(int64_t n, int chunks, const float* __restrict__ work, float* __restrict__ dw) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        float v = 0.f;
        for (int c = 0; c < chunks; ++c) v += work[(size_t)c * n + i];
        dw[i] = v;
    }
}

__global__ void flip_weights_kernel(int cin, int cout, const float* w, bf16_t* wb) {
    const int64_t n = (int64_t)cin * cout * 27;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        // i indexes wb[ci][t][co]
        const int co = (int)(i % cout);
        const int64_t r = i / cout;
        const int t = (int)(r % 27);
        const int ci = (int)(r / 27);
        wb[i] = f2bf(w[((int64_t)co * 27 + (26 - t)) * cin + ci]);
    }
}

__global__ void f32_to_bf16_kernel(int64_t n, const float* a, bf16_t* b) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        b[i] = f2bf(a[i]);
}

__global__ void relu_mask_kernel(int64_t n, const float* g, const bf16_t* act, bf16_t* d) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        d[i] = f2bf(bf2f(act[i]) > 0.f ? g[i] : 0.f);
}

// one thread per (coarse voxel, channel)
__global__ void avgpool_fwd_kernel(int n0, int n1, int n2, int c, const bf16_t* in, bf16_t* out) {
    const int m0 = n0 >> 1, m1 = n1 >> 1, m2 = n2 >> 1;
    const int64_t n = (int64_t)m0 * m1 * m2 * c;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int ch = (int)(i % c);
        const int64_t P = i / c;
        const int Z = (int)(P % m2);
        const int64_t Q = P / m2;
        const int Y = (int)(Q % m1);
        const int X = (int)(Q / m1);
        float s = 0.f;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int x = 2 * X + (k >> 2), y = 2 * Y + ((k >> 1) & 1), z = 2 * Z + (k & 1);
            s += bf2f(in[((int64_t)(x * n1 + y) * n2 + z) * c + ch]);
        }
        out[i] = f2bf(0.125f * s);
    }
}

// one thread per (fine voxel, channel)
__global__ void avgpool_bwd_kernel(int n0, int n1, int n2, int c, const float* dout, float* din) {
    const int m1 = n1 >> 1, m2 = n2 >> 1;
    const int64_t n = (int64_t)n0 * n1 * n2 * c;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int ch = (int)(i % c);
        const int64_t p = i / c;
        const int z = (int)(p % n2);
        const int64_t q = p / n2;
        const int y = (int)(q % n1);
        const int x = (int)(q / n1);
        din[i] += 0.125f * dout[((int64_t)((x >> 1) * m1 + (y >> 1)) * m2 + (z >> 1)) * c + ch];
    }
}

__global__ void upcat_fwd_kernel(int n0, int n1, int n2, int c0, int c1, const bf16_t* coarse, const bf16_t* fine,
                                 bf16_t* cat) {
    const int m1 = n1 >> 1, m2 = n2 >> 1;
    const int cc = c0 + c1;
    const int64_t n = (int64_t)n0 * n1 * n2 * cc;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int ch = (int)(i % cc);
        const int64_t p = i / cc;
        bf16_t v;
        if (ch < c0) {
            const int z = (int)(p % n2);
            const int64_t q = p / n2;
            const int y = (int)(q % n1);
            const int x = (int)(q / n1);
            v = coarse[((int64_t)((x >> 1) * m1 + (y >> 1)) * m2 + (z >> 1)) * c0 + ch];
        } else {
            v = fine[p * c1 + (ch - c0)];
        }
        cat[i] = v;
    }
}

// dcoarse: one thread per (coarse voxel, channel < c0); dfine: per (fine voxel, channel < c1)
__global__ void upcat_bwd_coarse_kernel(int n0, int n1, int n2, int c0, int c1, const float* dcat, float* dcoarse) {
    const int m0 = n0 >> 1, m1 = n1 >> 1, m2 = n2 >> 1;
    const int cc = c0 + c1;
    const int64_t n = (int64_t)m0 * m1 * m2 * c0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int ch = (int)(i % c0);
        const int64_t P = i / c0;
        const int Z = (int)(P % m2);
        const int64_t Q = P / m2;
        const int Y = (int)(Q % m1);
        const int X = (int)(Q / m1);
        float s = 0.f;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int x = 2 * X + (k >> 2), y = 2 * Y + ((k >> 1) & 1), z = 2 * Z + (k & 1);
            s += dcat[((int64_t)(x * n1 + y) * n2 + z) * cc + ch];
        }
        dcoarse[i] = s;
    }
}

__global__ void upcat_bwd_fine_kernel(int64_t nvox, int c0, int c1, const float* dcat, float* dfine) {
    const int cc = c0 + c1;
    const int64_t n = nvox * c1;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int ch = (int)(i % c1);
        const int64_t p = i / c1;
        dfine[i] = dcat[p * cc + c0 + ch];
    }
}

__global__ void adam_kernel(int64_t n, float* w, const float* g, float* m, float* v, float lr, float b1, float b2,
                            float eps, float c1, float c2, bf16_t* wbf) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
        w[i] = wi;
        wbf[i] = f2bf(wi);
    }
}

static unsigned grid_for(int64_t n, int bs = 256) {
    int64_t b = (n + bs - 1) / bs;
    if (b > 65536) b = 65536;
    return (unsigned)(b < 1 ? 1 : b);
}

hipError_t conv3d_fwd(const ConvShape& s, const bf16_t* in, const bf16_t* w, int relu, int out_kind, void* out,
                      int out_ld, int out_off, hipStream_t st) {
    if ((s.cin & 15) || (s.cout & 15) || s.cout < 16 || s.cout > 96 || s.cout == 80) return hipErrorInvalidValue;
    const int nvox = s.n0 * s.n1 * s.n2;
    if (nvox <= 0) return hipSuccess;
    const unsigned nb = (unsigned)((nvox + 63) / 64);
    const int NB = s.cout / 16;
#define STP_CONV_LAUNCH(NBv, OUTv)                                                                                  \
    hipLaunchKernelGGL((conv_fwd_kernel<NBv, OUTv>), dim3(nb), dim3(256), 0, st, s.n0, s.n1, s.n2, s.cin, in, w, relu, \
                       out, out_ld, out_off)
#define STP_CONV_NB(OUTv)                     \
    switch (NB) {                             \
        case 1: STP_CONV_LAUNCH(1, OUTv); break; \
        case 2: STP_CONV_LAUNCH(2, OUTv); break; \
        case 3: STP_CONV_LAUNCH(3, OUTv); break; \
        case 4: STP_CONV_LAUNCH(4, OUTv); break; \
        default: STP_CONV_LAUNCH(6, OUTv); break; \
    }
    if (out_kind == CONV_OUT_BF16) {
        STP_CONV_NB(CONV_OUT_BF16)
    } else if (out_kind == CONV_OUT_F32) {
        STP_CONV_NB(CONV_OUT_F32)
    } else {
        STP_CONV_NB(CONV_OUT_F32_SOA)
    }
#undef STP_CONV_NB
#undef STP_CONV_LAUNCH
    return hipGetLastError();
}

// voxel chunks of the weight gradient: about 8192 wavefronts in total, at
// least 512 voxels (a multiple of 32) each
static void bwd_weight_chunks(const ConvShape& s, int* chunks_out, int* vc_out) {
    const int nvox = s.n0 * s.n1 * s.n2;
    const int nblk = (s.cout / 16) * (27 * s.cin / 16);
    int chunks = 8192 / nblk;
    if (chunks < 1) chunks = 1;
    int vc = (nvox + chunks - 1) / chunks;
    vc = ((vc + 31) / 32) * 32;
    if (vc < 512) vc = 512;
    *chunks_out = (nvox + vc - 1) / vc;
    *vc_out = vc;
}

size_t conv3d_bwd_weight_work(const ConvShape& s) {
    if (s.n0 * s.n1 * s.n2 <= 0 || s.cin < 16 || s.cout < 16) return 0;
    int chunks = 0, vc = 0;
    bwd_weight_chunks(s, &chunks, &vc);
    return (size_t)chunks * s.cout * 27 * s.cin;
}

hipError_t conv3d_bwd_weight(const ConvShape& s, const bf16_t* in, const bf16_t* dout, float* dw, float* work,
                             hipStream_t st) {
    if ((s.cin & 15) || (s.cout & 15)) return hipErrorInvalidValue;
    const int nvox = s.n0 * s.n1 * s.n2;
    if (nvox <= 0) return hipSuccess;
    if (conv3d_bwd_weight_work(s) > CONV_BWD_WORK_MAX) return hipErrorInvalidValue;
    const int nblk = (s.cout / 16) * (27 * s.cin / 16);
    int chunks = 0, vc = 0;
    bwd_weight_chunks(s, &chunks, &vc);
    hipLaunchKernelGGL(conv_bwd_weight_kernel, dim3((nblk + 3) / 4, chunks), dim3(256), 0, st, s.n0, s.n1, s.n2, s.cin,
                       s.cout, in, dout, work, vc);
    const int64_t n = (int64_t)s.cout * 27 * s.cin;
    hipLaunchKernelGGL(sum_chunks_kernel, dim3(grid_for(n)), dim3(256), 0, st, n, chunks, work, dw);
    return hipGetLastError();
}

hipError_t conv3d_flip_weights(int cin, int cout, const float* w32, bf16_t* wb, hipStream_t st) {
    hipLaunchKernelGGL(flip_weights_kernel, dim3(grid_for((int64_t)cin * cout * 27)), dim3(256), 0, st, cin, cout, w32,
                       wb);
    return hipGetLastError();
}

hipError_t f32_to_bf16(int64_t n, const float* a, bf16_t* b, hipStream_t st) {
    hipLaunchKernelGGL(f32_to_bf16_kernel, dim3(grid_for(n)), dim3(256), 0, st, n, a, b);
    return hipGetLastError();
}

hipError_t relu_mask_bf16(int64_t n, const float* g, const bf16_t* act, bf16_t* d, hipStream_t st) {
    hipLaunchKernelGGL(relu_mask_kernel, dim3(grid_for(n)), dim3(256), 0, st, n, g, act, d);
    return hipGetLastError();
}

hipError_t avgpool2_fwd(int n0, int n1, int n2, int c, const bf16_t* in, bf16_t* out, hipStream_t st) {
    if ((n0 | n1 | n2) & 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(avgpool_fwd_kernel, dim3(grid_for((int64_t)n0 * n1 * n2 / 8 * c)), dim3(256), 0, st, n0, n1,
                       n2, c, in, out);
    return hipGetLastError();
}

hipError_t avgpool2_bwd(int n0, int n1, int n2, int c, const float* dout, float* din, hipStream_t st) {
    if ((n0 | n1 | n2) & 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(avgpool_bwd_kernel, dim3(grid_for((int64_t)n0 * n1 * n2 * c)), dim3(256), 0, st, n0, n1, n2, c,
                       dout, din);
    return hipGetLastError();
}

hipError_t upcat_fwd(int n0, int n1, int n2, int c0, int c1, const bf16_t* coarse, const bf16_t* fine, bf16_t* cat,
                     hipStream_t st) {
    if ((n0 | n1 | n2) & 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(upcat_fwd_kernel, dim3(grid_for((int64_t)n0 * n1 * n2 * (c0 + c1))), dim3(256), 0, st, n0, n1,
                       n2, c0, c1, coarse, fine, cat);
    return hipGetLastError();
}

hipError_t upcat_bwd(int n0, int n1, int n2, int c0, int c1, const float* dcat, float* dcoarse, float* dfine,
                     hipStream_t st) {
    if ((n0 | n1 | n2) & 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(upcat_bwd_coarse_kernel, dim3(grid_for((int64_t)n0 * n1 * n2 / 8 * c0)), dim3(256), 0, st, n0,
                       n1, n2, c0, c1, dcat, dcoarse);
    hipLaunchKernelGGL(upcat_bwd_fine_kernel, dim3(grid_for((int64_t)n0 * n1 * n2 * c1)), dim3(256), 0, st,
                       (int64_t)n0 * n1 * n2, c0, c1, dcat, dfine);
    return hipGetLastError();
}

hipError_t adam_step(int64_t n, float* w, const float* g, float* m, float* v, float lr, float b1, float b2, float eps,
                     float c1, float c2, bf16_t* wbf, hipStream_t st) {
    hipLaunchKernelGGL(adam_kernel, dim3(grid_for(n)), dim3(256), 0, st, n, w, g, m, v, lr, b1, b2, eps, c1, c2, wbf);
    return hipGetLastError();
}

}  // namespace stp
