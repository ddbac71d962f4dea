// Acoustic leapfrog FD step for gfx950 (MI355X).
//
// One launch advances every shot by one time step, in place on two buffers:
//   N = U + ((U - D*P) + C*L(U))        written over P
// which is the reference recurrence
//   next = 2*cur - prev + v^2 dt^2 * Lap(cur), then prev,cur *= D (sponge)
// (src/cpp/solver/wave_equation.cpp:262-297, 299-405) with the sponge folded
// in at read time: the reference damps `previous` after each step, here P is
// stored undamped and D is applied when it is read.  L is evaluated from
// neighbour-minus-centre differences so that its FP32 rounding error scales
// with |L| rather than with |U|.
//
// Work decomposition (x is the slowest axis, z contiguous):
//   * a workgroup owns a TZ x TY tile of the (y,z) plane and marches along x
//     over a chunk of LX planes;
//   * each thread owns VW consecutive z cells (one float4 when VW = 4) in TYR
//     rows and keeps a register queue of the 2R+1 planes x-R..x+R of its own
//     cells (x-derivative taps come from registers);
//   * the centre plane x is staged in LDS with an R-cell halo in y and z
//     (y- and z-derivative taps come from LDS); corners are never loaded
//     (star stencil); one barrier per plane;
//   * loads run PD planes ahead (queue front x+PD+R, P and C of x+PD) in
//     register rings whose slots are compile-time constants (the x loop is
//     unrolled by the ring length), so each plane waits only for its own
//     loads; the halo cells of a plane are fetched with its queue front,
//     R planes before it is the centre, when the neighbouring tiles fetch
//     the same cells as their own queue front, so they come from L2 (cut
//     the step's DRAM reads from 1.18x to 1.05x the algorithmic bytes), and
//     wait in that plane's LDS buffer (NB >= R+2 buffers);
//   * workgroup ids are remapped so each XCD gets a contiguous range of tiles
//     (neighbouring tiles share that XCD's L2 for their halos);
//   * a.rev marches each x chunk from its last plane down: the host
//     alternates it step by step, so a step starts on the planes the previous
//     one touched last, part of which the Infinity Cache still holds.
// Sponge: D = min(Dx(plane), Dyz(cell)) because the reference factor is a
// monotone function of min(distance) — Dyz is computed once per thread.
// Cerjan mode (damp_next, not in the reference) multiplies the result by D as
// well, N = D (U + (U - D*P) + C*L(U)): the reference form damps only the
// previous field, which makes smooth modes inside the sponge grow (root
// 1 + sqrt(1 - D) of its recurrence); this one decays (root D).
// Finite guards run only when the result is non-finite (a non-finite input
// always makes it so), which reproduces the reference's eager checks.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>
#include <utility>

#include "fd_kernels.h"

namespace stp {

// non-temporal access flags of the step (template parameter NTL)
enum { FD_NT_LOADS = 1, FD_NT_STORE = 2, FD_NT_COEF = 4 };

template <int VW> struct Vec;
typedef float f4n __attribute__((ext_vector_type(4)));
template <> struct Vec<1> {
    using T = float;
    using N = float;
    __device__ static inline T from_native(N v) { return v; }
    __device__ static inline N to_native(T v) { return v; }
    __device__ static inline float get(const T& v, int) { return v; }
    __device__ static inline void set(T& v, int, float x) { v = x; }
    __device__ static inline T zero() { return 0.f; }
};
template <> struct Vec<4> {
    using T = float4;
    using N = f4n;
    __device__ static inline T from_native(N v) { return make_float4(v.x, v.y, v.z, v.w); }
    __device__ static inline N to_native(T v) { return f4n{v.x, v.y, v.z, v.w}; }
    __device__ static inline float get(const T& v, int j) {
        return j == 0 ? v.x : (j == 1 ? v.y : (j == 2 ? v.z : v.w));
    }
    __device__ static inline void set(T& v, int j, float x) {
        if (j == 0) v.x = x; else if (j == 1) v.y = x; else if (j == 2) v.z = x; else v.w = x;
    }
    __device__ static inline T zero() { return make_float4(0.f, 0.f, 0.f, 0.f); }
};

__device__ __forceinline__ bool finitef(float x) { return __builtin_isfinite(x); }

// Sponge factor from a distance-to-interior-face d (reference formula).
__device__ __forceinline__ float sponge_of(int d, int sw, float sigma_over_sw2) {
    if (d < sw && d >= 0) {
        const float e = (float)(sw - d);
        return 1.0f - sigma_over_sw2 * e * e;
    }
    return 1.0f;
}

// distance component along one axis, as in apply_sponge_layer: min over the
// low face (i - hw) and high face (nh - hw - i - 1) when within the band.
__device__ __forceinline__ int sponge_dist(int i, int nh, int hw, int sw) {
    int d = sw;
    if (i < hw + sw) d = min(d, i - hw);
    if (i >= nh - hw - sw) d = min(d, nh - hw - i - 1);
    return d;
}

// One axis' pair of taps at distance k, in neighbour-minus-centre form: its
// rounding scales with the differences rather than with |u|.  (The plain
// pair sum for k >= 2, with their -2u share added once per cell, saves two
// VALU ops per pair but raised the error of a 2000-step 128^3 run against
// FP64 from 6e-6 to 1.3e-4.)  The fast and the guard path share it, so a
// cell's result does not depend on the path that computed it.
__device__ __forceinline__ float lap_term(float c, float p, float m, float u, float acc) {
    return fmaf(c, (p - u) + (m - u), acc);
}

__device__ __forceinline__ int xcd_remap(int b, int nb) {
    // hardware dispatches workgroup b to XCD (b % 8); give each XCD a
    // contiguous range of logical tiles.
    const int xcd = b & 7, k = b >> 3, q = nb >> 3, r = nb & 7;
    return xcd < r ? xcd * (q + 1) + k : r * (q + 1) + (xcd - r) * q + k;
}

__host__ __device__ constexpr int gcd_c(int a, int b) { return b == 0 ? a : gcd_c(b, a % b); }

// f(integral_constant<int, 0>) ... f(integral_constant<int, N-1>), unrolled
// by construction (register-array indices derived from I stay constant).
template <typename F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

template <int R, int VW, int TZ, int TY, int TYR, int LX, int NTL, int PD, int NB, int MODE>
__global__ void __launch_bounds__((TZ / VW) * (TY / TYR))
fd_step_kernel(const FdStepArgs a) {
    using V = Vec<VW>;
    using vec = typename V::T;
    constexpr int NTZ = TZ / VW;              // threads along z
    constexpr int NTY = TY / TYR;             // threads along y
    constexpr int NT = NTZ * NTY;
    constexpr int HV = (R + VW - 1) / VW;     // halo vectors per z side
    constexpr int LW = NTZ + 2 * HV;          // LDS row width (vectors)
    constexpr int LH = TY + 2 * R;            // LDS rows
    constexpr int NHALO = 2 * R * NTZ + 2 * HV * TY;
    constexpr int NH = (NHALO + NT - 1) / NT; // halo items per thread

    // ring lengths: NS register slots for P, C and halos (loads run PD
    // planes ahead), NSL LDS plane buffers, UF planes per unrolled pass
    constexpr int NS = PD < 2 ? PD + 1 : (PD < 4 ? 4 : 8);   // >= PD + 1, a power of two
    // NB > 0: halos fetched with the queue front (early), NB LDS plane buffers
    constexpr bool EH = NB > 0;
    static_assert(!EH || NB >= R + 2, "early halos need R + 2 LDS plane buffers");
    constexpr int NSL = EH ? NB : 2;
    constexpr int UM = NS / gcd_c(NS, NSL) * NSL;
    constexpr int UF = ((2 * R + 1 + PD + UM - 1) / UM) * UM;
    __shared__ vec lds[NSL][LH][LW];

    const int tid = threadIdx.x;
    const int tz = tid % NTZ, ty = tid / NTZ;

    // ---- decode workgroup -> (shot, x chunk, y tile, z tile)
    const int lb = xcd_remap(blockIdx.x, gridDim.x);
    int rem = lb;
    const int zt = rem % a.ntz; rem /= a.ntz;
    const int yt = rem % a.nty; rem /= a.nty;
    const int xc = rem % a.nxc; rem /= a.nxc;
    const int shot = rem;

    const int hw = a.hw, nyh = a.nyh, nzh = a.nzh, nxh = a.nxh;
    const int zb = hw + zt * TZ, yb = hw + yt * TY;
    const int x0 = hw + xc * LX;
    const int x1 = min(x0 + LX, nxh - hw);

    const float* __restrict__ U = a.ucur + (int64_t)shot * a.ncell;
    float* __restrict__ P = a.uprev + (int64_t)shot * a.ncell;
    const float* __restrict__ C = a.coef;

    // Offsets are 32-bit element offsets from the per-shot base (a shot's
    // field is < 2^32 elements, checked at plan creation).  Cells outside the
    // grid load from a clamped in-bounds address and are never stored, which
    // keeps the loads branch-free.
    const uint32_t plane = (uint32_t)nyh * (uint32_t)nzh;
    const uint32_t safe = (uint32_t)yb * nzh + (uint32_t)zb;

    // own cells: rvalid = interior (stored); positions inside the array but
    // outside the interior are halo cells whose (zero) values the valid
    // neighbours read, so they load from their true address; only positions
    // beyond the array use the clamped address (no valid cell reads those).
    const int z = zb + tz * VW;
    const bool zvalid = z < nzh - hw;
    const bool zinb = z + VW <= nzh;
    bool rvalid[TYR];
    uint32_t roff[TYR];   // load offset of own vector within a plane
#pragma unroll
    for (int r = 0; r < TYR; ++r) {
        const int y = yb + ty * TYR + r;
        rvalid[r] = zvalid && (y < nyh - hw);
        roff[r] = (zinb && y < nyh) ? (uint32_t)y * nzh + (uint32_t)z : safe;
    }

    // halo items: LDS slot (-1 = none) and in-plane offset
    int hslot[NH];
    uint32_t hoff[NH];
#pragma unroll
    for (int m = 0; m < NH; ++m) {
        const int i = tid + m * NT;
        int ly = 0, lz = 0;
        bool use = i < NHALO;
        if (use) {
            if (i < 2 * R * NTZ) {           // y halo rows (interior columns)
                const int row = i / NTZ, col = i % NTZ;
                ly = row < R ? row : row + TY;
                lz = HV + col;
            } else {                          // z halo columns (interior rows)
                const int j = i - 2 * R * NTZ;
                const int row = j / (2 * HV), col = j % (2 * HV);
                ly = R + row;
                lz = col < HV ? col : col + NTZ;
            }
        }
        hslot[m] = use ? ly * LW + lz : -1;
        const int gy = yb - R + ly;
        const int gz = zb + (lz - HV) * VW;
        const bool inb = use && gy >= 0 && gy < nyh && gz >= 0 && gz + VW <= nzh;
        hoff[m] = inb ? (uint32_t)gy * nzh + (uint32_t)gz : safe;
    }

    // sponge: Dyz per own cell (constant along x)
    const float sg = a.sigma_over_sw2;
    const int sw = a.sw;
    float dyz[TYR][VW];
#pragma unroll
    for (int r = 0; r < TYR; ++r) {
        const int y = yb + ty * TYR + r;
        const int dy = sponge_dist(y, nyh, hw, sw);
#pragma unroll
        for (int j = 0; j < VW; ++j) {
            const int dz = sponge_dist(z + j, nzh, hw, sw);
            dyz[r][j] = sponge_of(min(dy, dz), sw, sg);
        }
    }

    auto ldv = [](const float* base, uint32_t off) -> vec {
        return *reinterpret_cast<const vec*>(base + off);
    };
    // streamed-once operands (P, C): optionally non-temporal so they do not
    // evict the U planes whose halos neighbouring tiles re-read from L2
    auto ldnt = [](const float* base, uint32_t off) -> vec {
        using nv_t = typename Vec<VW>::N;
        return V::from_native(__builtin_nontemporal_load(reinterpret_cast<const nv_t*>(base + off)));
    };
    auto ldP = [&](const float* base, uint32_t off) -> vec {
        if constexpr ((NTL & FD_NT_LOADS) != 0) return ldnt(base, off);
        else return ldv(base, off);
    };
    auto ldC = [&](const float* base, uint32_t off) -> vec {
        if constexpr ((NTL & (FD_NT_LOADS | FD_NT_COEF)) != 0) return ldnt(base, off);
        else return ldv(base, off);
    };

    // ---- prefetch ring.  Loads run PD planes ahead of the plane being
    // computed (planes counted within the chunk, see px below): the U queue
    // holds planes i-R .. i+R+PD (UF slots, plane p in slot (p + R) % UF), P
    // and C are kept for planes i .. i+PD (NS slots, plane p in slot p % NS).
    // The x loop is unrolled by UF, a multiple of NS and NSL, so every slot
    // index is a compile-time constant:
    // no register moves, and the wait for a plane's loads covers only that
    // plane.  Halo rows and columns of a plane:
    //   EH: fetched together with the plane's queue front (R planes before
    //       it is the centre), when the neighbouring tiles fetch the same
    //       cells as their own queue fronts, so they come from L2; held in
    //       the plane's LDS buffer (NSL = NB >= R + 2 buffers, plane p in
    //       p % NSL) until it is the centre;
    //   otherwise: fetched PD planes before the plane is the centre, two LDS
    //       buffers.
    vec q[UF][TYR], pq[NS][TYR], cq[NS][TYR], hq[NS][NH];
    // Planes are numbered within the chunk (0 .. llast, and R past either
    // end for the queue); a.rev marches the chunk from its last plane down,
    // so a step starts on the planes the previous one touched last.  The x
    // taps enter the Laplacian as (u[x+k] - u) + (u[x-k] - u), whose value
    // does not depend on the order of the two terms, so both directions give
    // the same bits.
    const int llast = x1 - x0 - 1;
    const bool rev = a.rev != 0;
    auto px = [&](int i) -> int { return rev ? x1 - 1 - i : x0 + i; };
#pragma unroll
    for (int k = 0; k < 2 * R + PD; ++k) {
        const int pl = px(min(k - R, llast + R));   // beyond the chunk: in bounds, unused
        const uint32_t pb = (uint32_t)pl * plane;
#pragma unroll
        for (int r = 0; r < TYR; ++r) q[k][r] = ldv(U, pb + roff[r]);
    }
    constexpr int HA = EH ? R : 0;   // halo planes lead the centre by HA
    if constexpr (EH) {
        // halos of planes x0 .. x0+R-1 go straight to their LDS buffers
        // (ordered before the first reads by the first plane's barrier)
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const uint32_t pb = (uint32_t)px(min(k, llast)) * plane;
#pragma unroll
            for (int m = 0; m < NH; ++m)
                if (hslot[m] >= 0) (&lds[k % NSL][0][0])[hslot[m]] = ldv(U, pb + hoff[m]);
        }
    }
#pragma unroll
    for (int k = 0; k < PD; ++k) {
        const uint32_t pb = (uint32_t)px(min(k, llast)) * plane;
#pragma unroll
        for (int r = 0; r < TYR; ++r) {
            pq[k][r] = ldP(P, pb + roff[r]);
            cq[k][r] = ldC(C, pb + roff[r]);
        }
        const uint32_t hb = (uint32_t)px(min(HA + k, llast)) * plane;
#pragma unroll
        for (int m = 0; m < NH; ++m) hq[k][m] = ldv(U, hb + hoff[m]);
    }

    const float amp = a.amp;
    const bool damp_prev = a.damp_prev != 0;
    const bool damp_next = a.damp_next != 0;
    bool guard_hit = false;

    // one plane: chunk plane i = ib + ui (ui a compile-time constant, ib a
    // multiple of UF), grid plane x
    auto plane_step = [&](auto uic, const int ib) __attribute__((always_inline)) {
        constexpr int ui = decltype(uic)::value;
        const int i = ib + ui;
        const int x = px(i);
        // this plane's operands (compile-time slots: q[(ui + R + k) % UF] is chunk plane i + k)
        vec pc[TYR], cc[TYR];
#pragma unroll
        for (int r = 0; r < TYR; ++r) { pc[r] = pq[ui % NS][r]; cc[r] = cq[ui % NS][r]; }

        // stage plane i in LDS, and the halo of plane i + HA in its buffer
        vec(*L)[LW] = lds[ui % NSL];
#pragma unroll
        for (int r = 0; r < TYR; ++r) L[R + ty * TYR + r][HV + tz] = q[(ui + R) % UF][r];
#pragma unroll
        for (int m = 0; m < NH; ++m)
            if (hslot[m] >= 0) (&lds[(ui + HA) % NSL][0][0])[hslot[m]] = hq[ui % NS][m];

        // prefetch plane i+PD and the halo of plane i+PD+HA (planes past
        // the chunk re-read its last plane: in bounds, unused)
        {
            const int in = min(i + PD, llast);
            const uint32_t pb = (uint32_t)px(in) * plane;
            const uint32_t pf = (uint32_t)px(in + R) * plane;
#pragma unroll
            for (int r = 0; r < TYR; ++r) {
                q[(ui + 2 * R + PD) % UF][r] = ldv(U, pf + roff[r]);
                pq[(ui + PD) % NS][r] = ldP(P, pb + roff[r]);
                cq[(ui + PD) % NS][r] = ldC(C, pb + roff[r]);
            }
            const uint32_t hb = (uint32_t)px(min(i + PD + HA, llast)) * plane;
#pragma unroll
            for (int m = 0; m < NH; ++m) hq[(ui + PD) % NS][m] = ldv(U, hb + hoff[m]);
        }
        __syncthreads();

        const float dxs = sponge_of(sponge_dist(x, nxh, hw, sw), sw, sg);
        const uint32_t pb = (uint32_t)x * plane;
        vec res[TYR];
        vec lapv[TYR];
        bool bad = false;
#pragma unroll
        for (int r = 0; r < TYR; ++r) {
            const int lr = R + ty * TYR + r;
            // z window: HV vectors left, own, HV right
            vec zw[2 * HV + 1];
#pragma unroll
            for (int h = 0; h < HV; ++h) {
                zw[h] = L[lr][tz + h];
                zw[HV + 1 + h] = L[lr][HV + tz + 1 + h];
            }
            zw[HV] = q[(ui + R) % UF][r];
            // y column
            vec yc[2 * R + 1];
#pragma unroll
            for (int k = -R; k <= R; ++k) {
                const int rr = r + k;
                if (k == 0) yc[R] = q[(ui + R) % UF][r];
                else if (rr >= 0 && rr < TYR) yc[R + k] = q[(ui + R) % UF][rr];
                else yc[R + k] = L[lr + k][HV + tz];
            }
#pragma unroll
            for (int j = 0; j < VW; ++j) {
                const float u = V::get(q[(ui + R) % UF][r], j);
                float lap = 0.f;
#pragma unroll
                for (int k = 1; k <= R; ++k) {
                    const int zp = HV * VW + j + k, zm = HV * VW + j - k;
                    lap = lap_term(a.cx[k], V::get(q[(ui + R + k) % UF][r], j), V::get(q[(ui + R - k) % UF][r], j), u,
                                   lap);
                    lap = lap_term(a.cy[k], V::get(yc[R + k], j), V::get(yc[R - k], j), u, lap);
                    lap = lap_term(a.cz[k], V::get(zw[zp / VW], zp % VW), V::get(zw[zm / VW], zm % VW), u, lap);
                }
                const float p = V::get(pc[r], j);
                const float ds = fminf(dxs, dyz[r][j]);
                const float d = damp_prev ? ds : 1.0f;
                const float c = V::get(cc[r], j);
                float nv = u + (fmaf(-d, p, u) + c * lap);
                if (damp_next) nv *= ds;
                if constexpr (MODE == FD_MODE_SNAPSHOT_LAP) V::set(lapv[r], j, lap);
                bad |= !finitef(nv) && rvalid[r];
                // clamp to +-amp in one v_med3 (a non-finite nv is redone below)
                V::set(res[r], j, __builtin_amdgcn_fmed3f(nv, -amp, amp));
            }
        }
        if (bad) {
            // rare path, reference guard order (wave_equation.cpp:345-397): a
            // non-finite centre or previous value gives 0, a non-finite
            // neighbour keeps the centre value (both unclamped), a non-finite
            // result keeps the centre value (clamped).  The neighbours come
            // from the same registers and LDS plane as the fast path (no
            // global loads here, so the prefetch ring's waits stay exact).
            guard_hit = true;
#pragma unroll
            for (int r = 0; r < TYR; ++r) {
                if (!rvalid[r]) continue;
                const int lr = R + ty * TYR + r;
                vec zw[2 * HV + 1];
#pragma unroll
                for (int h = 0; h < HV; ++h) {
                    zw[h] = L[lr][tz + h];
                    zw[HV + 1 + h] = L[lr][HV + tz + 1 + h];
                }
                zw[HV] = q[(ui + R) % UF][r];
                vec yc[2 * R + 1];
#pragma unroll
                for (int k = -R; k <= R; ++k) {
                    const int rr = r + k;
                    if (k == 0) yc[R] = q[(ui + R) % UF][r];
                    else if (rr >= 0 && rr < TYR) yc[R + k] = q[(ui + R) % UF][rr];
                    else yc[R + k] = L[lr + k][HV + tz];
                }
#pragma unroll
                for (int j = 0; j < VW; ++j) {
                    const float u = V::get(q[(ui + R) % UF][r], j);
                    const float p = V::get(pc[r], j);
                    const float c = V::get(cc[r], j);
                    const float ds = fminf(dxs, dyz[r][j]);
                    const float d = damp_prev ? ds : 1.0f;
                    bool ok_n = true;
                    float lap = 0.f;
#pragma unroll
                    for (int k = 1; k <= R; ++k) {
                        const float xp = V::get(q[(ui + R + k) % UF][r], j);
                        const float xm = V::get(q[(ui + R - k) % UF][r], j);
                        const float yp = V::get(yc[R + k], j), ym = V::get(yc[R - k], j);
                        const int zpi = HV * VW + j + k, zmi = HV * VW + j - k;
                        const float zp = V::get(zw[zpi / VW], zpi % VW), zm = V::get(zw[zmi / VW], zmi % VW);
                        ok_n &= finitef(xp) && finitef(xm) && finitef(yp) && finitef(ym) && finitef(zp) &&
                                finitef(zm);
                        lap = lap_term(a.cx[k], xp, xm, u, lap);
                        lap = lap_term(a.cy[k], yp, ym, u, lap);
                        lap = lap_term(a.cz[k], zp, zm, u, lap);
                    }
                    // Cerjan mode multiplies the reference result by D
                    const float dn = damp_next ? ds : 1.0f;
                    float nv;
                    if (!(finitef(u) && finitef(p))) nv = 0.f;
                    else if (!ok_n) nv = dn * u;
                    else {
                        nv = dn * (u + (fmaf(-d, p, u) + c * lap));
                        nv = fminf(fmaxf(finitef(nv) ? nv : dn * u, -amp), amp);
                    }
                    V::set(res[r], j, nv);
                }
            }
        }
#pragma unroll
        for (int r = 0; r < TYR; ++r)
            if (rvalid[r]) {
                if constexpr ((NTL & FD_NT_STORE) != 0)
                    __builtin_nontemporal_store(V::to_native(res[r]),
                                                reinterpret_cast<typename Vec<VW>::N*>(P + pb + roff[r]));
                else *reinterpret_cast<vec*>(P + pb + roff[r]) = res[r];
                if constexpr (MODE == FD_MODE_SNAPSHOT || MODE == FD_MODE_SNAPSHOT_LAP) {
                    // read back only by the reverse pass of the segment: keep
                    // it out of the caches the stencil's planes live in
                    float* snap = a.snap_out + (int64_t)shot * a.ncell;
                    auto* dst = reinterpret_cast<typename Vec<VW>::N*>(snap + pb + roff[r]);
                    if constexpr (MODE == FD_MODE_SNAPSHOT_LAP)
                        __builtin_nontemporal_store(V::to_native(lapv[r]), dst);
                    else
                        __builtin_nontemporal_store(V::to_native(q[(ui + R) % UF][r]), dst);
                }
                if constexpr (MODE == FD_MODE_ACCUMULATE) {
                    // the snapshot is read once: non-temporal, like its store
                    const float* S = a.acc_src + (int64_t)shot * a.ncell;
                    const vec sv = V::from_native(__builtin_nontemporal_load(
                        reinterpret_cast<const typename Vec<VW>::N*>(S + pb + roff[r])));
                    const int y = yb + ty * TYR + r;
                    const int64_t ii = (int64_t)shot * a.nint +
                                       ((int64_t)(x - hw) * a.ny + (y - hw)) * a.nz + (z - hw);
                    vec* ao = reinterpret_cast<vec*>(a.acc_out + ii);
                    vec acc = *ao;
#pragma unroll
                    for (int j = 0; j < VW; ++j)
                        V::set(acc, j, fmaf(V::get(res[r], j), V::get(sv, j), V::get(acc, j)));
                    *ao = acc;
                    if (a.acc_ill) {
                        vec* io = reinterpret_cast<vec*>(a.acc_ill + ii);
                        vec ill = *io;
#pragma unroll
                        for (int j = 0; j < VW; ++j)
                            V::set(ill, j, fmaf(V::get(sv, j), V::get(sv, j), V::get(ill, j)));
                        *io = ill;
                    }
                }
            }
    };
    // whole ring lengths run unguarded, so every wait on a plane's loads
    // counts exactly the loads issued after them; the guarded tail (fewer
    // than UF planes) is a second copy
    const int xlen = llast + 1;
    int ib = 0;
    for (; ib + UF <= xlen; ib += UF)
        static_for<UF>([&](auto uic) __attribute__((always_inline)) { plane_step(uic, ib); });
    if (ib < xlen)
        static_for<UF>([&](auto uic) __attribute__((always_inline)) {
            if (ib + decltype(uic)::value < xlen) plane_step(uic, ib);   // uniform over the workgroup
        });

    if (guard_hit && a.step_status) atomicOr(a.step_status, STP_STEP_NAN_DETECTED_AND_REPAIRED_DEV);

    // ---- epilogue: receivers sample the new field, then the next source
    // sample is injected (the reference injects at the start of the next step).
    if (a.nrcv_pts == 0 && a.src_idx == nullptr) return;
    __syncthreads();
    const int tile = lb % (a.ntz * a.nty * a.nxc);
    if constexpr (MODE == FD_MODE_ACCUMULATE) {
        if (a.inj_traces) {
            // receiver injection into the new field (atomic: two receivers may
            // share a cell); the accumulated correlation gets the matching
            // correction so it sees the post-injection value
            const float* S = a.acc_src + (int64_t)shot * a.ncell;
            const int pbeg = a.blk_rcv_ptr[tile], pend = a.blk_rcv_ptr[tile + 1];
            for (int p = pbeg + tid; p < pend; p += NT) {
                const int64_t ci = a.rcv_cell[p];
                const int rid = a.rcv_id[p];
                float amt = a.inj_traces[((int64_t)shot * a.nrcv + rid) * a.inj_stride + a.inj_col];
                if (a.inj_scale_c) amt *= C[ci];
                const int cx = (int)(ci / (int64_t)plane), cy = (int)((ci / nzh) % nyh), cz = (int)(ci % nzh);
                if (damp_next)
                    amt *= sponge_of(min(sponge_dist(cx, nxh, hw, sw), min(sponge_dist(cy, nyh, hw, sw),
                                                                           sponge_dist(cz, nzh, hw, sw))),
                                     sw, sg);
                if (!finitef(amt)) continue;
                atomicAdd(&P[ci], amt);
                const int64_t ii = (int64_t)shot * a.nint + ((int64_t)(cx - hw) * a.ny + (cy - hw)) * a.nz + (cz - hw);
                atomicAdd(&a.acc_out[ii], amt * S[ci]);
            }
        }
        return;
    }
    if (a.traces) {
        const int pbeg = a.blk_rcv_ptr[tile], pend = a.blk_rcv_ptr[tile + 1];
        for (int p = pbeg + tid; p < pend; p += NT) {
            const int64_t ci = a.rcv_cell[p];
            const int rid = a.rcv_id[p];
            a.traces[((int64_t)shot * a.nrcv + rid) * a.trace_stride + a.t_col] = P[ci];
        }
    }
    if (a.src_idx && a.w_next) {
        __syncthreads();   // receivers read the pre-injection value
        if (tid == 0) {
            const int64_t s = a.src_idx[shot];
            const int sx = (int)(s / (int64_t)plane), sy = (int)((s / nzh) % nyh), sz = (int)(s % nzh);
            if (sx >= x0 && sx < x1 && sy >= yb && sy < yb + TY && sz >= zb && sz < zb + TZ) {
                const float w = a.w_next[(int64_t)shot * a.w_stride];
                if (finitef(w)) P[s] += w;
            }
        }
    }
}

// Injection of the first source sample of a propagate call (1 thread per shot).
__global__ void fd_inject_kernel(float* u, int64_t ncell, const int64_t* src_idx, const float* w,
                                 int64_t w_stride, int nshots) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nshots) return;
    const float v = w[(int64_t)s * w_stride];
    if (__builtin_isfinite(v)) u[(int64_t)s * ncell + src_idx[s]] += v;
}

// C = v^2 dt^2 * damp, with the reference's velocity substitution.
__global__ void fd_coef_kernel(const float* vel, float* coef, int64_t n, float dt2, float damp) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v = vel[i];
    if (!__builtin_isfinite(v) || v <= 0.f) v = 1500.f;
    coef[i] = v * v * dt2 * damp;
}

// max finite velocity (values are >= 0 after the filter, so the float bits
// order like unsigned ints and atomicMax on the bit pattern is exact).
__global__ void fd_vmax_kernel(const float* vel, int64_t n, unsigned int* out) {
    unsigned int best = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float v = vel[i];
        if (__builtin_isfinite(v) && v > 0.f) best = max(best, __float_as_uint(v));
    }
    for (int o = 32; o > 0; o >>= 1) best = max(best, (unsigned)__shfl_xor((int)best, o));
    if ((threadIdx.x & 63) == 0) atomicMax(out, best);
}

__global__ void fd_sponge_kernel(float* d, int nxh, int nyh, int nzh, int hw, int sw, float sg) {
    const int64_t n = (int64_t)nxh * nyh * nzh;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int ix = (int)(i / ((int64_t)nyh * nzh));
    const int iy = (int)((i / nzh) % nyh);
    const int iz = (int)(i % nzh);
    const int dd = min(sponge_dist(ix, nxh, hw, sw), min(sponge_dist(iy, nyh, hw, sw), sponge_dist(iz, nzh, hw, sw)));
    d[i] = sponge_of(dd, sw, sg);
}

// Streaming ceiling of the step's access pattern: out = a + b * c, three
// reads and one write of FP32 per cell (the step's 16 algorithmic bytes)
// with no stencil, one float4 per thread, optionally non-temporal.
template <bool NT>
__global__ void __launch_bounds__(256) stream3_kernel(const f4n* __restrict__ a, const f4n* __restrict__ b,
                                                      const f4n* __restrict__ c, f4n* __restrict__ o, int64_t n4) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n4) return;
    if constexpr (NT) {
        const f4n x = __builtin_nontemporal_load(a + i), y = __builtin_nontemporal_load(b + i),
                  z = __builtin_nontemporal_load(c + i);
        __builtin_nontemporal_store(x + y * z, o + i);
    } else {
        o[i] = a[i] + b[i] * c[i];
    }
}

hipError_t fd_launch_stream3(int64_t n, const float* a, const float* b, const float* c, float* o, bool nt,
                             hipStream_t st) {
    const int64_t n4 = n / 4;
    if (n4 == 0) return hipSuccess;
    const unsigned nb = (unsigned)((n4 + 255) / 256);
    if (nt)
        hipLaunchKernelGGL(stream3_kernel<true>, dim3(nb), dim3(256), 0, st, reinterpret_cast<const f4n*>(a),
                           reinterpret_cast<const f4n*>(b), reinterpret_cast<const f4n*>(c),
                           reinterpret_cast<f4n*>(o), n4);
    else
        hipLaunchKernelGGL(stream3_kernel<false>, dim3(nb), dim3(256), 0, st, reinterpret_cast<const f4n*>(a),
                           reinterpret_cast<const f4n*>(b), reinterpret_cast<const f4n*>(c),
                           reinterpret_cast<f4n*>(o), n4);
    return hipGetLastError();
}

// ------------------------------------------------------------------ launchers
template <int R, int VW, int TZ, int TY, int TYR, int LX, int NTL = 0, int PD = 1, int NB = 0>
static hipError_t launch_cfg(const FdStepArgs& a0, int nblocks, int mode, hipStream_t st) {
    constexpr int NT = (TZ / VW) * (TY / TYR);
    // the snapshot kind is a compile-time mode: a run-time choice between the
    // field and its Laplacian made the compiler spill the plane queue
    if (mode == FD_MODE_SNAPSHOT && a0.snap_lap)
        hipLaunchKernelGGL((fd_step_kernel<R, VW, TZ, TY, TYR, LX, NTL, PD, NB, FD_MODE_SNAPSHOT_LAP>), dim3(nblocks), dim3(NT),
                           0, st, a0);
    else if (mode == FD_MODE_SNAPSHOT)
        hipLaunchKernelGGL((fd_step_kernel<R, VW, TZ, TY, TYR, LX, NTL, PD, NB, FD_MODE_SNAPSHOT>), dim3(nblocks), dim3(NT), 0,
                           st, a0);
    else if (mode == FD_MODE_ACCUMULATE)
        hipLaunchKernelGGL((fd_step_kernel<R, VW, TZ, TY, TYR, LX, NTL, PD, NB, FD_MODE_ACCUMULATE>), dim3(nblocks), dim3(NT),
                           0, st, a0);
    else
        hipLaunchKernelGGL((fd_step_kernel<R, VW, TZ, TY, TYR, LX, NTL, PD, NB, FD_MODE_PLAIN>), dim3(nblocks), dim3(NT), 0, st,
                           a0);
    return hipGetLastError();
}

// Tile shapes (TZ, TY, TYR, LX, non-temporal flags, prefetch distance in
// planes, LDS plane buffers: 0 = halos fetched with the centre plane, else
// with the queue front) compiled for the float4 path; STP_FD_TILE=<index>
// selects one.  Entry 0 is the default of one-shot plans: 128 x 16 cells,
// one row per thread, 512 threads, x chunks of 256 planes (at 512^3, 256
// workgroups, one per CU), loads two planes ahead, halos fetched with the
// queue front, C (read once, never written) loaded non-temporally so it does
// not take L2 lines from the u^t planes the neighbouring tiles share.  Entry
// 1, the same without the non-temporal C, is the default of batched shots,
// which share C and re-read it from the caches (16 shots at 256^3: 4.37-4.49 s
// per gradient against 5.35 s with entry 0).  Entry 2 loads three planes
// ahead through 8 LDS buffers, whose 16-plane unrolled pass divides the
// 256-plane chunk (no guarded tail); entry 3 is the single-plane-ahead,
// late-halo kernel this replaced, kept as the reference point.  At 512^3,
// interleaved on one box: entry 0 381.3 us/step, 1 389.0, 3 ~411; on others
// 384.1-385.5 (entry 0) against 382.3-383.0 (entry 2), and 378.5-381.9
// against 381.4-386.2 in full propagations: within the spread of the boxes.
// These are upward-only sweeps; with the alternating sweeps the host uses,
// on one box: entries 0-3 378.2, 385.6, 377.8 and 404.4 us.
// Measured and not kept: three planes ahead with a guarded tail (390.3-390.8),
// P, C and N all non-temporal (388.5, against 389.0 for entry 1), late halos
// two planes ahead (398), 128 x 8 cells (422), 256 x 8 cells (397), shorter
// x chunks, which re-read 2R halo planes of u^t per chunk (64 planes:
// 451 us), and taller tiles (128 x 32, 256 x 16: 454-1171 us, or spills).
#define STP_FD_VARIANTS(X)                \
    X(0, 128, 16, 1, 256, 4, 2, 6)        \
    X(1, 128, 16, 1, 256, 0, 2, 6)        \
    X(2, 128, 16, 1, 256, 4, 3, 8)        \
    X(3, 128, 16, 1, 256, 0, 1, 0)

int fd_num_variants() { return 4; }

// Fixed tiles of the other paths (must match fd_launch_step below).
#define STP_FD_FIXED(X)                 \
    X(4, 1, 64, 8, 2, 64)               \
    X(1, 1, 64, 8, 2, 64)               \
    X(1, 4, 64, 16, 2, 64)

void fd_tiling(int R, int VW, int variant, int* tz, int* ty, int* lx) {
#define STP_FIXED_TILE(RR, VV, TZ, TY, TYR, LX) \
    if (R == RR && VW == VV) { *tz = TZ; *ty = TY; *lx = LX; return; }
    STP_FD_FIXED(STP_FIXED_TILE)
#undef STP_FIXED_TILE
#define STP_TILE_CASE(id, TZ, TY, TYR, LX, NTL, PD, NB) \
    if (variant == id) { *tz = TZ; *ty = TY; *lx = LX; return; }
    STP_FD_VARIANTS(STP_TILE_CASE)
#undef STP_TILE_CASE
    *tz = 128; *ty = 8; *lx = 64;
}

hipError_t fd_launch_step(const FdStepArgs& a, int R, int VW, int variant, int mode, hipStream_t st) {
    const int nblocks = a.ntz * a.nty * a.nxc * a.nshots;
#define STP_FIXED_LAUNCH(RR, VV, TZ, TY, TYR, LX) \
    if (R == RR && VW == VV) return launch_cfg<RR, VV, TZ, TY, TYR, LX>(a, nblocks, mode, st);
    STP_FD_FIXED(STP_FIXED_LAUNCH)
#undef STP_FIXED_LAUNCH
    if (R != 4 || VW != 4) return hipErrorInvalidValue;
#define STP_LAUNCH_CASE(id, TZ, TY, TYR, LX, NTL, PD, NB) \
    if (variant == id) return launch_cfg<4, 4, TZ, TY, TYR, LX, NTL, PD, NB>(a, nblocks, mode, st);
    STP_FD_VARIANTS(STP_LAUNCH_CASE)
#undef STP_LAUNCH_CASE
    return hipErrorInvalidValue;
}

hipError_t fd_launch_inject(float* u, int64_t ncell, const int64_t* src_idx, const float* w, int64_t w_stride,
                            int nshots, hipStream_t st) {
    hipLaunchKernelGGL(fd_inject_kernel, dim3((nshots + 63) / 64), dim3(64), 0, st, u, ncell, src_idx, w, w_stride,
                       nshots);
    return hipGetLastError();
}

hipError_t fd_launch_coef(const float* vel, float* coef, int64_t n, float dt2, float damp, hipStream_t st) {
    hipLaunchKernelGGL(fd_coef_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, vel, coef, n, dt2, damp);
    return hipGetLastError();
}

hipError_t fd_launch_vmax(const float* vel, int64_t n, unsigned int* out, hipStream_t st) {
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(fd_vmax_kernel, dim3((unsigned)blocks), dim3(256), 0, st, vel, n, out);
    return hipGetLastError();
}

hipError_t fd_launch_sponge(float* d, int nxh, int nyh, int nzh, int hw, int sw, float sg, hipStream_t st) {
    const int64_t n = (int64_t)nxh * nyh * nzh;
    hipLaunchKernelGGL(fd_sponge_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d, nxh, nyh, nzh, hw,
                       sw, sg);
    return hipGetLastError();
}

}  // namespace stp
