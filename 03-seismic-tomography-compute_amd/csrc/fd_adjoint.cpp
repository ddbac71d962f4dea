// Reverse-time pass of the FD propagator: RTM imaging and the adjoint-state
// gradient (declarations: include/seismic_hip.h).
//
// Forward pass: nt steps from rest, checkpointing (u^t0 pre-injection,
// previous) every K steps.  Reverse pass, segment by segment from the last:
// restore the checkpoint, recompute the segment's K steps writing one
// snapshot per step (the step's input field u~^t, or its Laplacian L u~^t),
// then run the adjoint recursion backwards through the segment:
//   phi^k = 2 phi^{k+1} - D phi^{k+2} + C L phi^{k+1} + C g^k,
// which is the forward step applied to (phi^{k+1}, phi^{k+2}) followed by the
// receiver injection C(x_r) e_r^{k-1}; each backward step accumulates
// acc += phi^k * snapshot^{k-1} (and acc_ill += snapshot^2).
// With snapshots of L u~ the accumulator is dJ/dC * C for J = 1/2 |d - d_obs|^2
// (e = residual traces): dJ/dv = 2 acc / v.  With snapshots of u~ it is the
// zero-lag cross-correlation image of source and receiver wavefields.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "capi_common.h"
#include "fd_plan.h"
#include "seismic_hip.h"

using stp::fail;
using stp::S;

namespace {

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
};

// e = d - o (FP32) and per-block FP64 partial sums of e^2
__global__ void residual_misfit_kernel(int64_t n, const float* d, const float* o, float* e, double* part) {
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float r = d[i] - o[i];
        e[i] = r;
        s += (double)r * (double)r;
    }
    __shared__ double red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// run `nsteps` forward steps; *u/*v are swapped as the steps proceed.
// snap (mode snapshot): base of nsteps slots of nshots*ncell floats.
int forward_steps(stp_fd_plan* p, float** u, float** v, const float* coef, int nsteps, int prev_damped,
                  const float* wavelet, int nw, int t0, float* traces, int trace_stride, float* snap, int snap_lap,
                  hipStream_t st) {
    if (nsteps <= 0) return STP_OK;
    stp::FdStepArgs a;
    stp::fd_fill_args(p, &a);
    a.coef = coef;
    a.nrcv_pts = traces ? p->nrcv_pts : 0;
    a.traces = traces;
    a.trace_stride = trace_stride;
    const bool inject = wavelet && nw > 0 && p->d_src;
    a.src_idx = inject ? p->d_src : nullptr;
    a.w_stride = nw;
    a.snap_lap = snap_lap;
    if (inject && t0 < nw)
        STP_HIP(stp::fd_launch_inject(*u, p->ncell, p->d_src, wavelet + t0, nw, p->cfg.nshots, st));
    const int64_t slot = (int64_t)p->cfg.nshots * p->ncell;
    for (int s = 0; s < nsteps; ++s) {
        const int t = t0 + s;
        a.ucur = *u;
        a.uprev = *v;
        a.damp_prev = (s == 0) ? !prev_damped : 1;
        a.t_col = s;
        a.w_next = (inject && s + 1 < nsteps && t + 1 < nw) ? wavelet + (t + 1) : nullptr;
        a.snap_out = snap ? snap + s * slot : nullptr;
        a.rev = p->alt_sweep ? t & 1 : 0;   // alternate sweeps (same bits either way)
        STP_HIP(stp::fd_launch_step(a, p->R, p->VW, p->variant,
                                    snap ? stp::FD_MODE_SNAPSHOT : stp::FD_MODE_PLAIN, st));
        std::swap(*u, *v);
    }
    return STP_OK;
}

}  // namespace

extern "C" {

int stp_fd_adjoint(stp_fd_plan* p, const float* coef, const float* wavelet, int32_t nw, int32_t nt,
                   const float* obs, const float* inj_traces, int32_t inj_scale_c, int32_t snap_lap,
                   int32_t checkpoint_interval, float* acc_out, float* acc_ill, float* traces_out, double* misfit,
                   void* stream) {
    if (!p || !coef || nt < 1 || nw < 0 || (!acc_out && !obs))
        return fail(STP_ERR_INVALID, "stp_fd_adjoint: bad argument");
    const bool misfit_only = !acc_out;   // forward pass and misfit, no reverse pass
    if ((obs || inj_traces || traces_out) && p->nrcv == 0)
        return fail(STP_ERR_STATE, "stp_fd_adjoint: traces given but no receivers set");
    if (obs && inj_traces) return fail(STP_ERR_INVALID, "stp_fd_adjoint: give either obs or inj_traces");
    if (p->VW == 4 && (p->cfg.nz % 4) != 0) return fail(STP_ERR_INVALID, "stp_fd_adjoint: unsupported shape");
    hipStream_t st = S(stream);
    const int nshots = p->cfg.nshots;
    int K = checkpoint_interval > 0 ? checkpoint_interval : (int)std::ceil(std::sqrt((double)nt));
    K = std::min(K, nt);
    const int nck = (nt + K - 1) / K;
    const int64_t field = (int64_t)nshots * p->ncell;          // floats per multi-shot field
    const int64_t nint = p->cfg.nx * p->cfg.ny * p->cfg.nz;
    DevBuf work, ckpt, seg, tr, res, part;
    const int64_t ntr = (int64_t)nshots * p->nrcv * nt;
    if (obs) {
        if (!traces_out) {
            STP_HIP(hipMalloc(&tr.p, sizeof(float) * ntr));
            traces_out = static_cast<float*>(tr.p);
        }
        STP_HIP(hipMalloc(&res.p, sizeof(float) * ntr));
        STP_HIP(hipMalloc(&part.p, sizeof(double) * 1024));
    }
    STP_HIP(hipMalloc(&work.p, sizeof(float) * field * (misfit_only ? 2 : 4)));
    if (!misfit_only) {
        STP_HIP(hipMalloc(&ckpt.p, sizeof(float) * field * 2 * nck));
        STP_HIP(hipMalloc(&seg.p, sizeof(float) * field * K));
    }
    float* w0 = static_cast<float*>(work.p);
    float* fu = w0;
    float* fv = w0 + field;
    float* au = w0 + 2 * field;
    float* av = w0 + 3 * field;
    float* ck = static_cast<float*>(ckpt.p);
    float* sg = static_cast<float*>(seg.p);

    // forward pass with checkpoints
    STP_HIP(hipMemsetAsync(w0, 0, sizeof(float) * field * 2, st));
    for (int c = 0; c < nck; ++c) {
        const int t0 = c * K, len = std::min(K, nt - t0);
        if (!misfit_only) {
            STP_HIP(hipMemcpyAsync(ck + (2 * c) * field, fu, sizeof(float) * field, hipMemcpyDeviceToDevice, st));
            STP_HIP(hipMemcpyAsync(ck + (2 * c + 1) * field, fv, sizeof(float) * field, hipMemcpyDeviceToDevice,
                                   st));
        }
        int rc = forward_steps(p, &fu, &fv, coef, len, 0, wavelet, nw, t0, traces_out ? traces_out + t0 : nullptr, nt,
                               nullptr, 0, st);
        if (rc != STP_OK) return rc;
    }

    // residual traces and misfit
    if (obs) {
        const unsigned nb = (unsigned)std::min<int64_t>(1024, (ntr + 255) / 256);
        hipLaunchKernelGGL(residual_misfit_kernel, dim3(nb), dim3(256), 0, st, ntr, traces_out, obs,
                           static_cast<float*>(res.p), static_cast<double*>(part.p));
        STP_HIP(hipGetLastError());
        std::vector<double> hp(nb);
        STP_HIP(hipMemcpyAsync(hp.data(), part.p, sizeof(double) * nb, hipMemcpyDeviceToHost, st));
        STP_HIP(hipStreamSynchronize(st));
        double j = 0.0;
        for (double v : hp) j += v;
        if (misfit) *misfit = 0.5 * j;
        inj_traces = static_cast<const float*>(res.p);
    }

    if (misfit_only) return STP_OK;   // the misfit read above synchronised the stream

    // reverse pass
    STP_HIP(hipMemsetAsync(au, 0, sizeof(float) * field * 2, st));
    STP_HIP(hipMemsetAsync(acc_out, 0, sizeof(float) * nshots * nint, st));
    if (acc_ill) STP_HIP(hipMemsetAsync(acc_ill, 0, sizeof(float) * nshots * nint, st));
    stp::FdStepArgs a;
    stp::fd_fill_args(p, &a);
    a.coef = coef;
    a.nrcv_pts = inj_traces ? p->nrcv_pts : 0;
    a.acc_out = acc_out;
    a.acc_ill = acc_ill;
    a.inj_traces = inj_traces;
    a.inj_stride = nt;
    a.inj_scale_c = inj_scale_c;
    a.damp_prev = 1;
    for (int c = nck - 1; c >= 0; --c) {
        const int t0 = c * K, len = std::min(K, nt - t0);
        fu = w0;
        fv = w0 + field;
        STP_HIP(hipMemcpyAsync(fu, ck + (2 * c) * field, sizeof(float) * field, hipMemcpyDeviceToDevice, st));
        STP_HIP(hipMemcpyAsync(fv, ck + (2 * c + 1) * field, sizeof(float) * field, hipMemcpyDeviceToDevice, st));
        int rc = forward_steps(p, &fu, &fv, coef, len, 0, wavelet, nw, t0, nullptr, 0, sg, snap_lap, st);
        if (rc != STP_OK) return rc;
        for (int t = t0 + len - 1; t >= t0; --t) {
            a.ucur = au;
            a.uprev = av;
            a.acc_src = sg + (int64_t)(t - t0) * field;
            a.inj_col = t;
            a.rev = p->alt_sweep ? t & 1 : 0;
            STP_HIP(stp::fd_launch_step(a, p->R, p->VW, p->variant, stp::FD_MODE_ACCUMULATE, st));
            std::swap(au, av);
        }
    }
    STP_HIP(hipStreamSynchronize(st));
    return STP_OK;
}

}  // extern "C"
