// Two leapfrog steps per launch (temporal blocking) for the 8th-order
// stencil on gfx950.  Per cell and pair of steps it moves 20 bytes (read U,
// P, C; write V, W) instead of the 32 of two single steps.
//
//   step 1: V = U + ((U - D1*P) + C*L(U))      V~ = V + w[t+1] at the source
//   step 2: W = V~ + ((V~ - D*U) + C*L(V~))
//
// U = u~^t (source already injected), P = u^{t-1} (undamped unless the
// call says otherwise: D1 = D or 1), D the sponge factor.  The outputs go to
// two buffers that no workgroup of this launch reads, so workgroups that
// recompute a neighbour's cells (below) never see updated values:
//   Vout = V (pre-injection; fd_step2_post adds w[t+1] after the receivers
//          sampled it, making Vout the next call's undamped previous field)
//   Wout = W (the next current field; fd_step2_post samples it and adds
//          w[t+2])
//
// Work decomposition: a workgroup owns a TZ x TY tile of the (y,z) plane
// and a chunk of LX planes along x.  Step 2 at plane x needs V on planes
// x-R..x+R and on the tile widened by R in y and z, so each workgroup runs
// step 1 on the widened tile ((TZ+2R) x (TY+2R), recomputing a ring that
// its neighbours also compute) R planes ahead of step 2.  One thread per
// float4 column of the widened tile:
//   * a register queue of U on planes xa-R..xa+R (x taps of step 1, and U
//     at xb = xa-R as step 2's previous field),
//   * the U plane xa (widened by another R) staged in LDS for the y and z
//     taps of step 1, double buffered, one barrier per plane;
//   * a ring of the last 2R+1 V planes of the widened tile in LDS: step 2
//     takes its x taps from its own column there and its y and z taps from
//     plane xb, which every thread wrote R planes (and barriers) earlier.
// The arithmetic of each step is written exactly as in fd_step_kernel, so
// the two kernels give bit-identical fields.  Every global load is
// predicated on its plane and in-plane position lying inside the array;
// cells outside the interior are never computed (their V is 0) and never
// stored.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fd_kernels.h"

namespace stp {

namespace {

__device__ __forceinline__ bool fin(float x) { return __builtin_isfinite(x); }

__device__ __forceinline__ float sp_of(int d, int sw, float s) {
    if (d < sw && d >= 0) {
        const float e = (float)(sw - d);
        return 1.0f - s * e * e;
    }
    return 1.0f;
}

__device__ __forceinline__ int sp_dist(int i, int nh, int hw, int sw) {
    int d = sw;
    if (i < hw + sw) d = min(d, i - hw);
    if (i >= nh - hw - sw) d = min(d, nh - hw - i - 1);
    return d;
}

__device__ __forceinline__ int xcd_map(int b, int nb) {
    const int xcd = b & 7, k = b >> 3, q = nb >> 3, r = nb & 7;
    return xcd < r ? xcd * (q + 1) + k : r * (q + 1) + (xcd - r) * q + k;
}

__device__ __forceinline__ float g4(const float4& v, int j) {
    return j == 0 ? v.x : (j == 1 ? v.y : (j == 2 ? v.z : v.w));
}
__device__ __forceinline__ void s4(float4& v, int j, float x) {
    if (j == 0) v.x = x; else if (j == 1) v.y = x; else if (j == 2) v.z = x; else v.w = x;
}

// L(u) of the four cells of a float4 column: x taps from the plane queue
// q[0..8] (centre q[4]), z taps from the centre vector and its left / right
// neighbours, y taps from ycol(k), k = -4..4 (k != 0).  Per cell the terms
// are added in the order of fd_step_kernel: for k = 1..4, x, y, z.
template <typename YCol>
__device__ __forceinline__ float4 laplacian(const float4* q, const float4 zl, const float4 zr, YCol ycol,
                                            const FdStep2Args& a) {
    const float4 c = q[4];
    float lap[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 1; k <= 4; ++k) {
        const float4 xp = q[4 + k], xm = q[4 - k], yp = ycol(k), ym = ycol(-k);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float u = g4(c, j);
            const float xs = (g4(xp, j) - u) + (g4(xm, j) - u);
            const float ys = (g4(yp, j) - u) + (g4(ym, j) - u);
            const int zp = j + k, zm = j - k;
            const float up = zp < 4 ? g4(c, zp) : g4(zr, zp - 4);
            const float um = zm >= 0 ? g4(c, zm) : g4(zl, zm + 4);
            const float zs = (up - u) + (um - u);
            lap[j] = fmaf(a.cx[k], xs, lap[j]);
            lap[j] = fmaf(a.cy[k], ys, lap[j]);
            lap[j] = fmaf(a.cz[k], zs, lap[j]);
        }
    }
    return make_float4(lap[0], lap[1], lap[2], lap[3]);
}

// new = u + ((u - d p) + c L), clamped to +-amp, with the reference's guards
// (wave_equation.cpp:345-397): a non-finite centre or previous value gives 0,
// a non-finite neighbour (hence a non-finite L) keeps the centre value, a
// non-finite result keeps the centre value.  `bad` is set when a guard fired.
__device__ __forceinline__ float4 update(const float4 u4, const float4 p4, const float4 c4, const float4 l4,
                                         const float* d, float amp, bool& bad) {
    float4 r;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float u = g4(u4, j), p = g4(p4, j), c = g4(c4, j), lap = g4(l4, j);
        const float nv = u + (fmaf(-d[j], p, u) + c * lap);
        float o = fminf(fmaxf(nv, -amp), amp);
        if (!fin(nv)) {
            bad = true;
            o = !(fin(u) && fin(p)) ? 0.f : (!fin(lap) ? u : fminf(fmaxf(u, -amp), amp));
        }
        s4(r, j, o);
    }
    return r;
}

}  // namespace

// TZ, TY: owned tile (cells); LX: planes per chunk.  R = 4, float4 columns.
template <int TZ, int TY, int LX>
__global__ void __launch_bounds__((TZ / 4 + 2) * (TY + 8))
fd_step2_kernel(const FdStep2Args a) {
    constexpr int R = 4;
    constexpr int EZ = TZ / 4 + 2;          // widened tile, float4 columns
    constexpr int EY = TY + 2 * R;          // widened tile, rows
    constexpr int NT = EZ * EY;
    constexpr int AZ = EZ + 2;              // U plane in LDS: widened tile + R
    constexpr int AY = EY + 2 * R;
    constexpr int NHALO = 2 * R * EZ + 2 * EY;   // star halo of the U plane
    constexpr int NV = 2 * R + 1;                // V planes kept
    static_assert(NHALO <= NT, "one halo item per thread at most");

    __shared__ float4 LA[2][AY][AZ];
    __shared__ float4 LV[NV][EY][EZ];

    const int tid = threadIdx.x;
    const int tz = tid % EZ, ty = tid / EZ;

    const int lb = xcd_map(blockIdx.x, gridDim.x);
    int rem = lb;
    const int zt = rem % a.ntz; rem /= a.ntz;
    const int yt = rem % a.nty; rem /= a.nty;
    const int xc = rem % a.nxc; rem /= a.nxc;
    const int shot = rem;

    const int hw = a.hw, nxh = a.nxh, nyh = a.nyh, nzh = a.nzh;
    const int zb = hw + zt * TZ, yb = hw + yt * TY;
    const int x0 = hw + xc * LX;
    const int x1 = min(x0 + LX, nxh - hw);
    const uint32_t plane = (uint32_t)nyh * (uint32_t)nzh;

    const float* __restrict__ U = a.u + (int64_t)shot * a.ncell;
    const float* __restrict__ P = a.p + (int64_t)shot * a.ncell;
    const float* __restrict__ C = a.coef;
    float* __restrict__ Vo = a.vout + (int64_t)shot * a.ncell;
    float* __restrict__ Wo = a.wout + (int64_t)shot * a.ncell;

    // this thread's column: cells (y, z..z+3)
    const int y = yb - R + ty;
    const int z = zb - 4 + 4 * tz;
    const bool col_in = y >= 0 && y < nyh && z >= 0 && z + 4 <= nzh;            // inside the array
    const bool col_int = y >= hw && y < nyh - hw && z >= hw && z + 4 <= nzh - hw;  // interior cells
    const uint32_t coff = col_in ? (uint32_t)y * nzh + (uint32_t)z : 0u;
    const bool own = tz >= 1 && tz <= TZ / 4 && ty >= R && ty < R + TY && col_int;

    // halo item of the U plane (star: y rows above/below the widened tile,
    // z columns left/right of it)
    int hy = 0, hz = 0;
    bool huse = false;
    if (tid < 2 * R * EZ) {
        const int row = tid / EZ, c = tid % EZ;
        hy = row < R ? row : row + EY;
        hz = 1 + c;
        huse = true;
    } else if (tid < NHALO) {
        const int j = tid - 2 * R * EZ;
        hy = R + j / 2;
        hz = (j & 1) ? AZ - 1 : 0;
        huse = true;
    }
    const int gy = yb - 2 * R + hy, gz = zb - 8 + 4 * hz;
    const bool h_in = huse && gy >= 0 && gy < nyh && gz >= 0 && gz + 4 <= nzh;
    const uint32_t hoff = h_in ? (uint32_t)gy * nzh + (uint32_t)gz : 0u;

    const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
    auto ld = [&](const float* base, int px, bool ok, uint32_t off) -> float4 {
        if (ok && px >= 0 && px < nxh) return *reinterpret_cast<const float4*>(base + (uint32_t)px * plane + off);
        return zero4;
    };

    // sponge: Dyz of the column (constant along x)
    const float sg = a.sigma_over_sw2;
    const int sw = a.sw;
    float dyz[4];
    {
        const int dy = sp_dist(y, nyh, hw, sw);
#pragma unroll
        for (int j = 0; j < 4; ++j) dyz[j] = sp_of(min(dy, sp_dist(z + j, nzh, hw, sw)), sw, sg);
    }

    // source cell of this shot inside this column (lane), for the mid-pair injection
    int src_x = -1, src_lane = -1;
    float w_mid = 0.f;
    if (a.src_idx && a.w_mid) {
        const int64_t s = a.src_idx[shot];
        const int sx = (int)(s / (int64_t)plane), sy = (int)((s / nzh) % nyh), sz = (int)(s % nzh);
        const float w = a.w_mid[(int64_t)shot * a.w_stride];
        if (sy == y && sz >= z && sz < z + 4 && fin(w)) {
            src_x = sx;
            src_lane = sz - z;
            w_mid = w;
        }
    }

    const float amp = a.amp;
    const bool damp1 = a.damp_prev != 0;
    bool bad1 = false, bad2 = false;

    // U queue (C at the step-2 plane is read again: it was fetched R planes
    // earlier by this thread and is still in L2)
    float4 uq[2 * R + 1];
    const int xs = x0 - R;            // first step-1 plane
#pragma unroll
    for (int k = 0; k <= 2 * R; ++k) uq[k] = zero4;
#pragma unroll
    for (int k = 1; k <= 2 * R; ++k) uq[k] = ld(U, xs - R - 1 + k, col_in, coff);   // planes xs-R .. xs+R-1
    float4 un = ld(U, xs + R, col_in, coff);
    float4 pn = ld(P, xs, col_in, coff);
    float4 cn = ld(C, xs, col_in, coff);
    float4 c2n = ld(C, xs - R, own, coff);
    float4 hn = ld(U, xs, h_in, hoff);

    const int niter = (x1 - x0) + 2 * R;
    for (int it = 0; it < niter; ++it) {
        const int xa = xs + it;       // step-1 plane
        const int xb = xa - R;        // step-2 plane
#pragma unroll
        for (int k = 0; k < 2 * R; ++k) uq[k] = uq[k + 1];
        uq[2 * R] = un;
        const float4 pc = pn, c1 = cn, c2 = c2n;

        float4(*A)[AZ] = LA[it & 1];
        A[R + ty][1 + tz] = uq[R];
        if (huse) A[hy][hz] = hn;

        // prefetch for the next plane (reads beyond the array return 0)
        un = ld(U, xa + 1 + R, col_in, coff);
        pn = ld(P, xa + 1, col_in, coff);
        cn = ld(C, xa + 1, col_in, coff);
        c2n = ld(C, xb + 1, own, coff);
        hn = ld(U, xa + 1, h_in, hoff);
        __syncthreads();

        // ---------------- step 1 at plane xa on the widened tile
        float4 vn = zero4;
        const bool do1 = col_int && xa >= hw && xa < nxh - hw;
        if (do1) {
            const float dx1 = sp_of(sp_dist(xa, nxh, hw, sw), sw, sg);
            const int lr = R + ty, lc = 1 + tz;
            const float4 lap = laplacian(uq, A[lr][lc - 1], A[lr][lc + 1], [&](int k) { return A[lr + k][lc]; }, a);
            float dd[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) dd[j] = damp1 ? fminf(dx1, dyz[j]) : 1.0f;
            bool bad = false;
            vn = update(uq[R], pc, c1, lap, dd, amp, bad);
            if (own) {
                bad1 |= bad;
                *reinterpret_cast<float4*>(Vo + (uint32_t)xa * plane + coff) = vn;
            }
            if (xa == src_x) s4(vn, src_lane, g4(vn, src_lane) + w_mid);
        }
        // slot of plane x in the V ring (xa >= x0 - R >= 0)
        LV[xa % NV][ty][tz] = vn;

        // ---------------- step 2 at plane xb on the owned tile
        if (own && it >= 2 * R && xb < x1) {
            const float dx2 = sp_of(sp_dist(xb, nxh, hw, sw), sw, sg);
            float4 vx[2 * R + 1];       // own column, planes xb-R..xb+R
#pragma unroll
            for (int k = -R; k <= R; ++k) vx[R + k] = LV[(xb + k) % NV][ty][tz];
            const int sb = xb % NV;
            const float4 lap =
                laplacian(vx, LV[sb][ty][tz - 1], LV[sb][ty][tz + 1], [&](int k) { return LV[sb][ty + k][tz]; }, a);
            float dd[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) dd[j] = fminf(dx2, dyz[j]);
            const float4 wn = update(vx[R], uq[0], c2, lap, dd, amp, bad2);
            *reinterpret_cast<float4*>(Wo + (uint32_t)xb * plane + coff) = wn;
        }
    }
    if (bad1 && a.status1) atomicOr(a.status1, STP_STEP_NAN_DETECTED_AND_REPAIRED_DEV);
    if (bad2 && a.status2) atomicOr(a.status2, STP_STEP_NAN_DETECTED_AND_REPAIRED_DEV);
}

// After a pair: receivers sample V (step-1 column) and W (step-2 column),
// then the source gets w[t+1] in V and w[t+2] in W.  One block per shot.
__global__ void fd_step2_post_kernel(const FdStep2Args a) {
    const int shot = blockIdx.x;
    float* V = a.vout + (int64_t)shot * a.ncell;
    float* W = a.wout + (int64_t)shot * a.ncell;
    if (a.traces) {
        for (int r = threadIdx.x; r < a.nrcv; r += blockDim.x) {
            const int64_t ci = a.rcv_cell_plain[r];
            float* tr = a.traces + ((int64_t)shot * a.nrcv + r) * a.trace_stride + a.t_col;
            tr[0] = V[ci];
            tr[1] = W[ci];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0 && a.src_idx) {
        const int64_t s = a.src_idx[shot];
        if (a.w_mid) {
            const float w = a.w_mid[(int64_t)shot * a.w_stride];
            if (fin(w)) V[s] += w;
        }
        if (a.w_next) {
            const float w = a.w_next[(int64_t)shot * a.w_stride];
            if (fin(w)) W[s] += w;
        }
    }
}

// Tiles (owned TZ x TY cells, LX planes per chunk), selectable through
// STP_FD2_TILE=<index>; each must keep (TZ/4+2)*(TY+8) <= 1024 threads and
// its LDS ((TY+16)*(TZ/4+4)*2 + 9*(TY+8)*(TZ/4+2) float4) within 160 KB.
#define STP_FD2_VARIANTS(X) \
    X(0, 64, 32, 128)       \
    X(1, 64, 16, 128)       \
    X(2, 32, 32, 128)       \
    X(3, 128, 16, 128)      \
    X(4, 64, 32, 256)       \
    X(5, 32, 16, 128)

int fd_step2_num_variants() { return 6; }

void fd_step2_tiling(int variant, int* tz, int* ty, int* lx) {
#define STP_FD2_TILE(id, TZ, TY, LX) \
    if (variant == id) { *tz = TZ; *ty = TY; *lx = LX; return; }
    STP_FD2_VARIANTS(STP_FD2_TILE)
#undef STP_FD2_TILE
    *tz = 64; *ty = 32; *lx = 128;
}

hipError_t fd_launch_step2(const FdStep2Args& a, int variant, hipStream_t st) {
    const int nblocks = a.ntz * a.nty * a.nxc * a.nshots;
    hipError_t e = hipErrorInvalidValue;
#define STP_FD2_LAUNCH(id, TZ, TY, LX)                                                                        \
    if (variant == id) {                                                                                      \
        hipLaunchKernelGGL((fd_step2_kernel<TZ, TY, LX>), dim3(nblocks), dim3((TZ / 4 + 2) * (TY + 8)), 0, st, a); \
        e = hipGetLastError();                                                                                \
    }
    STP_FD2_VARIANTS(STP_FD2_LAUNCH)
#undef STP_FD2_LAUNCH
    if (e != hipSuccess) return e;
    if (a.traces || (a.src_idx && (a.w_mid || a.w_next))) {
        hipLaunchKernelGGL(fd_step2_post_kernel, dim3(a.nshots), dim3(256), 0, st, a);
        e = hipGetLastError();
    }
    return e;
}

}  // namespace stp
