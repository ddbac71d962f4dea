// C ABI of the seismic compute library (declarations: include/seismic_hip.h).
// Handles own only the small device tables they build (receiver lists,
// source indices, scratch words); wavefields, coefficients, traces and
// solver vectors are caller-owned device memory.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "fd_kernels.h"
#include "health_kernels.h"
#include "capi_common.h"
#include "fd_plan.h"
#include "seismic_hip.h"

namespace stp {
std::string& last_error() {
    thread_local std::string e;
    return e;
}
}  // namespace stp

namespace {
using stp::fail;
using stp::hip_fail;
using stp::S;

int fd_weights(int order, double* a) {
    if (order == 2) { a[0] = -2.0; a[1] = 1.0; return 1; }
    if (order == 8) {
        a[0] = -205.0 / 72.0; a[1] = 8.0 / 5.0; a[2] = -1.0 / 5.0; a[3] = 8.0 / 315.0; a[4] = -1.0 / 560.0;
        return 4;
    }
    return -1;
}

}  // namespace

namespace stp {
// common launch arguments of a plan (pointers and per-step fields left zero)
void fd_fill_args(const stp_fd_plan* p, FdStepArgs* a) {
    const stp_fd_config& c = p->cfg;
    std::memset(a, 0, sizeof *a);
    a->ncell = p->ncell;
    a->nxh = p->nxh; a->nyh = p->nyh; a->nzh = p->nzh; a->hw = c.halo_width;
    std::memcpy(a->cx, p->cx, sizeof a->cx);
    std::memcpy(a->cy, p->cy, sizeof a->cy);
    std::memcpy(a->cz, p->cz, sizeof a->cz);
    a->sw = c.sponge_width;
    a->sigma_over_sw2 = c.sponge_width > 0 ? (float)(c.sponge_damping / ((double)c.sponge_width * c.sponge_width)) : 0.f;
    a->amp = (float)c.amplitude_limit;
    a->damp_next = p->sponge_mode == STP_SPONGE_CERJAN;
    a->ntz = p->ntz; a->nty = p->nty; a->nxc = p->nxc; a->nshots = c.nshots;
    a->nrcv = p->nrcv;
    a->blk_rcv_ptr = p->d_blk_ptr;
    a->rcv_cell = p->d_rcv_cell;
    a->rcv_id = p->d_rcv_id;
    a->nx = (int)c.nx; a->ny = (int)c.ny; a->nz = (int)c.nz;
    a->nint = c.nx * c.ny * c.nz;
}
}  // namespace stp

extern "C" {

const char* stp_last_error(void) { return stp::last_error().c_str(); }
int stp_version(void) { return 100; }

int stp_device_alloc(void** p, int64_t bytes) {
    if (!p || bytes < 0) return fail(STP_ERR_INVALID, "stp_device_alloc: bad argument");
    *p = nullptr;
    if (bytes == 0) return STP_OK;
    hipError_t e = hipMalloc(p, (size_t)bytes);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc");
    return STP_OK;
}
int stp_device_free(void* p) {
    if (p) STP_HIP(hipFree(p));
    return STP_OK;
}
int stp_memcpy_h2d(void* d, const void* h, int64_t bytes, void* st) {
    if (bytes > 0) STP_HIP(hipMemcpyAsync(d, h, (size_t)bytes, hipMemcpyHostToDevice, S(st)));
    if (bytes > 0) STP_HIP(hipStreamSynchronize(S(st)));
    return STP_OK;
}
int stp_memcpy_d2h(void* h, const void* d, int64_t bytes, void* st) {
    if (bytes > 0) STP_HIP(hipMemcpyAsync(h, d, (size_t)bytes, hipMemcpyDeviceToHost, S(st)));
    if (bytes > 0) STP_HIP(hipStreamSynchronize(S(st)));
    return STP_OK;
}
int stp_memcpy_d2d(void* d, const void* s, int64_t bytes, void* st) {
    if (bytes > 0) STP_HIP(hipMemcpyAsync(d, s, (size_t)bytes, hipMemcpyDeviceToDevice, S(st)));
    return STP_OK;
}
int stp_memset_d(void* d, int v, int64_t bytes, void* st) {
    if (bytes > 0) STP_HIP(hipMemsetAsync(d, v, (size_t)bytes, S(st)));
    return STP_OK;
}
int stp_stream_sync(void* st) {
    STP_HIP(hipStreamSynchronize(S(st)));
    return STP_OK;
}
int stp_device_count(int* n) {
    STP_HIP(hipGetDeviceCount(n));
    return STP_OK;
}

// ------------------------------------------------------------------ FD plan
int stp_fd_plan_create(const stp_fd_config* c, stp_fd_plan** out) {
    if (!c || !out) return fail(STP_ERR_INVALID, "stp_fd_plan_create: null argument");
    *out = nullptr;
    double a[5];
    const int R = fd_weights(c->order, a);
    if (R < 0) return fail(STP_ERR_INVALID, "stp_fd_plan_create: order must be 2 or 8");
    if (c->nx < 1 || c->ny < 1 || c->nz < 1) return fail(STP_ERR_INVALID, "stp_fd_plan_create: empty grid");
    if (c->halo_width < R) return fail(STP_ERR_INVALID, "stp_fd_plan_create: halo_width < order/2");
    if (!(c->dx > 0 && c->dy > 0 && c->dz > 0 && c->dt > 0))
        return fail(STP_ERR_INVALID, "stp_fd_plan_create: spacings and dt must be positive");
    if (c->nshots < 1) return fail(STP_ERR_INVALID, "stp_fd_plan_create: nshots < 1");
    const int64_t nxh = c->nx + 2 * c->halo_width, nyh = c->ny + 2 * c->halo_width, nzh = c->nz + 2 * c->halo_width;
    if (nxh > (1 << 30) || nyh > (1 << 30) || nzh > (1 << 30) || nxh * nyh * nzh >= ((int64_t)1 << 32))
        return fail(STP_ERR_INVALID, "stp_fd_plan_create: a shot's field must have fewer than 2^32 cells");
    stp_fd_plan* p = new (std::nothrow) stp_fd_plan();
    if (!p) return fail(STP_ERR_ALLOC, "stp_fd_plan_create: out of host memory");
    p->cfg = *c;
    p->R = R;
    p->VW = (c->halo_width % 4 == 0 && c->nz % 4 == 0) ? 4 : 1;
    // One shot reads C once per step, so it is loaded non-temporally
    // (variant 0); shots of a batch share C and read it again from the
    // caches, so they keep it there (variant 1, the same tiling).
    p->variant = c->nshots == 1 ? 0 : 1;
    // tile-shape override for tuning: STP_FD_TILE=<variant index>
    if (const char* ev = std::getenv("STP_FD_TILE")) {
        const int v = std::atoi(ev);
        if (v >= 0 && v < stp::fd_num_variants()) p->variant = v;
    }
    if (const char* ev = std::getenv("STP_FD_SWEEP")) p->alt_sweep = std::strcmp(ev, "up") != 0;
    stp::fd_tiling(R, p->VW, p->variant, &p->tz, &p->ty, &p->lx);
    p->nxh = (int)nxh; p->nyh = (int)nyh; p->nzh = (int)nzh;
    p->ncell = nxh * nyh * nzh;
    p->ntz = (int)((c->nz + p->tz - 1) / p->tz);
    p->nty = (int)((c->ny + p->ty - 1) / p->ty);
    p->nxc = (int)((c->nx + p->lx - 1) / p->lx);
    const int64_t nblk = (int64_t)p->ntz * p->nty * p->nxc * c->nshots;
    if (nblk > 0x7fffffff) { delete p; return fail(STP_ERR_INVALID, "stp_fd_plan_create: too many tiles"); }
    for (int k = 0; k <= R; ++k) {
        p->cx[k] = (float)(a[k] / (c->dx * c->dx));
        p->cy[k] = (float)(a[k] / (c->dy * c->dy));
        p->cz[k] = (float)(a[k] / (c->dz * c->dz));
    }
    hipError_t e = hipMalloc(&p->d_word, 16);
    if (e != hipSuccess) { delete p; return hip_fail(e, "hipMalloc"); }
    *out = p;
    return STP_OK;
}

int stp_fd_plan_destroy(stp_fd_plan* p) {
    if (!p) return STP_OK;
    (void)hipFree(p->d_blk_ptr);
    (void)hipFree(p->d_rcv_cell);
    (void)hipFree(p->d_rcv_id);
    (void)hipFree(p->d_src);
    (void)hipFree(p->d_word);
    delete p;
    return STP_OK;
}

static bool interior_coords(const stp_fd_plan* p, int64_t c, int64_t* ix, int64_t* iy, int64_t* iz) {
    if (c < 0 || c >= p->ncell) return false;
    const int64_t plane = (int64_t)p->nyh * p->nzh;
    *ix = c / plane; *iy = (c / p->nzh) % p->nyh; *iz = c % p->nzh;
    const int hw = p->cfg.halo_width;
    return *ix >= hw && *ix < p->nxh - hw && *iy >= hw && *iy < p->nyh - hw && *iz >= hw && *iz < p->nzh - hw;
}

int stp_fd_set_receivers(stp_fd_plan* p, const int64_t* idx, int32_t n) {
    if (!p || n < 0 || (n > 0 && !idx)) return fail(STP_ERR_INVALID, "stp_fd_set_receivers: bad argument");
    const int ntile = p->ntz * p->nty * p->nxc;
    std::vector<int> cnt(ntile + 1, 0);
    std::vector<int> tile(n);
    const int hw = p->cfg.halo_width;
    for (int r = 0; r < n; ++r) {
        int64_t ix, iy, iz;
        if (!interior_coords(p, idx[r], &ix, &iy, &iz))
            return fail(STP_ERR_INVALID, "stp_fd_set_receivers: receiver " + std::to_string(r) + " is not an interior cell");
        const int t = (int)(((ix - hw) / p->lx * p->nty + (iy - hw) / p->ty) * p->ntz + (iz - hw) / p->tz);
        tile[r] = t;
        cnt[t + 1]++;
    }
    for (int t = 0; t < ntile; ++t) cnt[t + 1] += cnt[t];
    std::vector<int64_t> cell(n);
    std::vector<int> id(n);
    std::vector<int> pos(cnt.begin(), cnt.end() - 1);
    for (int r = 0; r < n; ++r) {
        const int k = pos[tile[r]]++;
        cell[k] = idx[r];
        id[k] = r;
    }
    (void)hipFree(p->d_blk_ptr); (void)hipFree(p->d_rcv_cell); (void)hipFree(p->d_rcv_id);
    p->d_blk_ptr = nullptr; p->d_rcv_cell = nullptr; p->d_rcv_id = nullptr;
    p->nrcv = 0; p->nrcv_pts = 0;
    if (n == 0) return STP_OK;
    STP_HIP(hipMalloc(&p->d_blk_ptr, sizeof(int) * (ntile + 1)));
    STP_HIP(hipMalloc(&p->d_rcv_cell, sizeof(int64_t) * n));
    STP_HIP(hipMalloc(&p->d_rcv_id, sizeof(int) * n));
    STP_HIP(hipMemcpy(p->d_blk_ptr, cnt.data(), sizeof(int) * (ntile + 1), hipMemcpyHostToDevice));
    STP_HIP(hipMemcpy(p->d_rcv_cell, cell.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice));
    STP_HIP(hipMemcpy(p->d_rcv_id, id.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    p->nrcv = n;
    p->nrcv_pts = n;
    return STP_OK;
}

int stp_fd_set_sources(stp_fd_plan* p, const int64_t* idx, int32_t n) {
    if (!p || !idx || n != p->cfg.nshots) return fail(STP_ERR_INVALID, "stp_fd_set_sources: need one index per shot");
    for (int s = 0; s < n; ++s) {
        int64_t ix, iy, iz;
        if (!interior_coords(p, idx[s], &ix, &iy, &iz))
            return fail(STP_ERR_INVALID, "stp_fd_set_sources: source is not an interior cell");
    }
    (void)hipFree(p->d_src);
    p->d_src = nullptr;
    STP_HIP(hipMalloc(&p->d_src, sizeof(int64_t) * n));
    STP_HIP(hipMemcpy(p->d_src, idx, sizeof(int64_t) * n, hipMemcpyHostToDevice));
    return STP_OK;
}

int stp_fd_coefficients(stp_fd_plan* p, const float* vel, float* coef, double* vmax_host, void* st) {
    if (!p || !vel || !coef) return fail(STP_ERR_INVALID, "stp_fd_coefficients: null argument");
    STP_HIP(hipMemsetAsync(p->d_word, 0, 4, S(st)));
    STP_HIP(stp::fd_launch_vmax(vel, p->ncell, p->d_word, S(st)));
    unsigned int bits = 0;
    STP_HIP(hipMemcpyAsync(&bits, p->d_word, 4, hipMemcpyDeviceToHost, S(st)));
    STP_HIP(hipStreamSynchronize(S(st)));
    float vmaxf;
    std::memcpy(&vmaxf, &bits, 4);
    const double vmax = vmaxf;
    const stp_fd_config& c = p->cfg;
    // wave_equation.cpp:316-331, 383-387: sub-cycling damp when the per-axis
    // CFL number exceeds 1
    const double cfl = std::max({vmax * c.dt / c.dx, vmax * c.dt / c.dy, vmax * c.dt / c.dz});
    double damp = 1.0;
    if (cfl > 1.0) {
        const double ratio = vmax * c.dt / std::min({c.dx, c.dy, c.dz});
        const double f = std::min(1.0, c.cfl_safety_factor / ratio);
        damp = f * f;
    }
    STP_HIP(stp::fd_launch_coef(vel, coef, p->ncell, (float)(c.dt * c.dt), (float)damp, S(st)));
    if (vmax_host) *vmax_host = vmax;
    return STP_OK;
}

int stp_fd_set_sponge_mode(stp_fd_plan* p, int32_t mode) {
    if (!p) return fail(STP_ERR_INVALID, "stp_fd_set_sponge_mode: null plan");
    if (mode != STP_SPONGE_REFERENCE && mode != STP_SPONGE_CERJAN)
        return fail(STP_ERR_INVALID, "stp_fd_set_sponge_mode: unknown mode");
    p->sponge_mode = mode;
    return STP_OK;
}

int stp_fd_verify_cfl(const stp_fd_plan* p, double vmax, int32_t* ok) {
    if (!p || !ok) return fail(STP_ERR_INVALID, "stp_fd_verify_cfl: null argument");
    const stp_fd_config& c = p->cfg;
    const double cm = std::max({vmax * c.dt / c.dx, vmax * c.dt / c.dy, vmax * c.dt / c.dz});
    int good = cm <= c.cfl_safety_factor;
    if (good && c.order != 2) {
        double a[5];
        const int R = fd_weights(c.order, a);
        double s = a[0];
        for (int k = 1; k <= R; ++k) s += 2.0 * a[k] * ((k & 1) ? -1.0 : 1.0);
        const double lam = std::fabs(s) * (1.0 / (c.dx * c.dx) + 1.0 / (c.dy * c.dy) + 1.0 / (c.dz * c.dz));
        good = vmax * vmax * c.dt * c.dt * lam <= 4.0;
    }
    *ok = good;
    return STP_OK;
}

int stp_fd_propagate(stp_fd_plan* p, float* cur, float* prev, const float* coef, int32_t nsteps, int32_t prev_damped,
                     const float* wavelet, int32_t nw, int32_t t_offset, float* traces, int32_t trace_stride,
                     int32_t* status, void* st) {
    if (!p || !cur || !prev || !coef || nsteps < 0 || t_offset < 0)
        return fail(STP_ERR_INVALID, "stp_fd_propagate: bad argument");
    if (traces && p->nrcv == 0) return fail(STP_ERR_STATE, "stp_fd_propagate: traces given but no receivers set");
    if (traces && trace_stride < nsteps) return fail(STP_ERR_INVALID, "stp_fd_propagate: trace_stride < nsteps");
    if (wavelet && nw > 0 && !p->d_src) return fail(STP_ERR_STATE, "stp_fd_propagate: wavelet given but no sources set");
    if (nsteps == 0) return STP_OK;
    const stp_fd_config& c = p->cfg;
    stp::FdStepArgs a;
    stp::fd_fill_args(p, &a);
    a.coef = coef;
    a.nrcv_pts = traces ? p->nrcv_pts : 0;
    a.traces = traces;
    a.trace_stride = trace_stride;
    const bool inject = wavelet && nw > 0 && p->d_src;
    a.src_idx = inject ? p->d_src : nullptr;
    a.w_stride = nw;
    if (inject && t_offset < nw)
        STP_HIP(stp::fd_launch_inject(cur, p->ncell, p->d_src, wavelet + t_offset, nw, c.nshots, S(st)));
    float* u = cur;
    float* v = prev;
    auto single = [&](int s) -> int {
        const int t = t_offset + s;
        a.ucur = u;
        a.uprev = v;
        a.damp_prev = (s == 0) ? !prev_damped : 1;
        a.t_col = s;
        a.w_next = (inject && s + 1 < nsteps && t + 1 < nw) ? wavelet + (t + 1) : nullptr;
        a.step_status = status ? status + s : nullptr;
        a.rev = p->alt_sweep ? t & 1 : 0;   // alternate sweeps: start on the planes the last step left in the caches
        STP_HIP(stp::fd_launch_step(a, p->R, p->VW, p->variant, stp::FD_MODE_PLAIN, S(st)));
        std::swap(u, v);
        return STP_OK;
    };
    int rc;
    int s = 0;
    for (; s < nsteps; ++s)
        if ((rc = single(s)) != STP_OK) return rc;
    return STP_OK;
}

int stp_stream_3r1w_f32(int64_t n, const float* a, const float* b, const float* c, float* out, int32_t nontemporal,
                        void* st) {
    if (n < 0 || n % 4 != 0 || (n > 0 && (!a || !b || !c || !out)))
        return fail(STP_ERR_INVALID, "stp_stream_3r1w_f32: n must be a non-negative multiple of 4");
    const uintptr_t al = reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b) |
                         reinterpret_cast<uintptr_t>(c) | reinterpret_cast<uintptr_t>(out);
    if (al % 16 != 0) return fail(STP_ERR_INVALID, "stp_stream_3r1w_f32: arrays must be 16-byte aligned");
    STP_HIP(stp::fd_launch_stream3(n, a, b, c, out, nontemporal != 0, S(st)));
    return STP_OK;
}

int stp_fd_sponge_factors(const stp_fd_plan* p, float* d, void* st) {
    if (!p || !d) return fail(STP_ERR_INVALID, "stp_fd_sponge_factors: null argument");
    const stp_fd_config& c = p->cfg;
    const float sg = c.sponge_width > 0 ? (float)(c.sponge_damping / ((double)c.sponge_width * c.sponge_width)) : 0.f;
    STP_HIP(stp::fd_launch_sponge(d, p->nxh, p->nyh, p->nzh, c.halo_width, c.sponge_width, sg, S(st)));
    return STP_OK;
}

// ------------------------------------------------------------------ field diagnostics
int stp_field_energy_f32(const float* u, int64_t n, double* out, void* st) {
    if (!out || n < 0 || (n > 0 && !u)) return fail(STP_ERR_INVALID, "stp_field_energy_f32: bad argument");
    STP_HIP(stp::field_energy_f32(u, n, out, S(st)));
    return STP_OK;
}

int stp_field_interior_f64(const float* u, int64_t nxh, int64_t nyh, int64_t nzh, int32_t hw, double* out,
                           void* st) {
    if (!u || !out || hw < 0 || nxh < 2 * hw || nyh < 2 * hw || nzh < 2 * hw)
        return fail(STP_ERR_INVALID, "stp_field_interior_f64: bad argument");
    STP_HIP(stp::field_interior_f64(u, nxh, nyh, nzh, hw, out, S(st)));
    return STP_OK;
}

// ------------------------------------------------------------------ health
int stp_count_nonfinite_f64(const double* d, int64_t n, int64_t* out, void* st) {
    if (!out || n < 0 || (n > 0 && !d)) return fail(STP_ERR_INVALID, "stp_count_nonfinite_f64: bad argument");
    *out = 0;
    if (n == 0) return STP_OK;
    STP_HIP(stp::health_count(d, n, out, S(st)));
    return STP_OK;
}
int stp_clamp_field_f64(double* d, int64_t n, double lo, double hi, void* st) {
    if (n < 0 || (n > 0 && !d)) return fail(STP_ERR_INVALID, "stp_clamp_field_f64: bad argument");
    if (n == 0) return STP_OK;
    STP_HIP(stp::health_clamp(d, n, lo, hi, S(st)));
    return STP_OK;
}
int stp_repair_field_nan_f64(double* d, int64_t count, int64_t nx, int64_t ny, int64_t nz, int32_t hw, double repl,
                             void* st) {
    if (count < 0 || (count > 0 && !d) || hw < 0) return fail(STP_ERR_INVALID, "stp_repair_field_nan_f64: bad argument");
    if (count == 0) return STP_OK;
    STP_HIP(stp::health_repair(d, count, nx, ny, nz, hw, repl, S(st)));
    return STP_OK;
}

}  // extern "C"
