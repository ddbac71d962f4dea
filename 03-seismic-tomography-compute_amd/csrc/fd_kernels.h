// Internal (C++) interface between the C ABI layer and the FD kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define STP_STEP_NAN_DETECTED_AND_REPAIRED_DEV 1

namespace stp {

// Per-launch arguments of the leapfrog step (passed by value, lives in SGPRs).
struct FdStepArgs {
    const float* ucur;        // current field, all shots
    float* uprev;             // previous field in, next field out (in place)
    const float* coef;        // C = v^2 dt^2 (shared by all shots)
    int64_t ncell;            // nxh*nyh*nzh (per-shot stride)
    int nxh, nyh, nzh, hw;
    float cx[5], cy[5], cz[5];  // weights a_k / h^2 per axis (index 1..R used)
    int sw;                   // sponge width
    float sigma_over_sw2;     // sponge_damping / sw^2
    float amp;                // amplitude clamp
    int damp_prev;            // apply the sponge factor to the previous field
    int damp_next;            // Cerjan sponge: the new field is multiplied by D too
    int ntz, nty, nxc, nshots;  // tiles along z, y; x chunks; shots
    // receivers (per-tile CSR lists)
    int nrcv_pts;             // total list entries (0 = no receivers)
    int nrcv;                 // receivers per shot
    const int* blk_rcv_ptr;   // [ntz*nty*nxc + 1]
    const int64_t* rcv_cell;  // with-halo cell index
    const int* rcv_id;        // receiver number
    float* traces;            // [shot][nrcv][trace_stride] or null
    int64_t trace_stride;
    int t_col;                // column written this step
    // source injection of the next step's sample (null = none)
    const int64_t* src_idx;   // [nshots]
    const float* w_next;      // &wavelet[t+1] of shot 0; shot s at s*w_stride
    int64_t w_stride;
    int* step_status;         // OR-ed with STP_STEP_NAN_DETECTED_AND_REPAIRED
    int rev;                  // march each x chunk from its last plane to its first
    // mode 1 (forward recompute): per-shot with-halo snapshot of the step's
    // input field (snap_lap = 0) or of its Laplacian L(u) (snap_lap = 1)
    float* snap_out;
    int snap_lap;
    // mode 2 (backward pass): acc_out += N * acc_src, acc_ill += acc_src^2
    // (interior-indexed, per shot), then the receivers inject
    // inj_traces[(shot*nrcv + r)*inj_stride + inj_col] (times C at the
    // receiver when inj_scale_c, times D there in Cerjan mode) into N
    const float* acc_src;
    float* acc_out;
    float* acc_ill;
    int64_t nint;             // interior cells per shot
    int nx, ny, nz;           // interior extents
    const float* inj_traces;
    int64_t inj_stride;
    int inj_col;
    int inj_scale_c;
};

// FD_MODE_SNAPSHOT is what callers pass (snap_lap picks the kind); the
// kernel is compiled separately for the field (1) and Laplacian (3) kinds.
enum { FD_MODE_PLAIN = 0, FD_MODE_SNAPSHOT = 1, FD_MODE_ACCUMULATE = 2, FD_MODE_SNAPSHOT_LAP = 3 };


int fd_num_variants();
hipError_t fd_launch_stream3(int64_t n, const float* a, const float* b, const float* c, float* o, bool nt,
                             hipStream_t st);
void fd_tiling(int R, int VW, int variant, int* tz, int* ty, int* lx);
hipError_t fd_launch_step(const FdStepArgs& a, int R, int VW, int variant, int mode, hipStream_t st);
hipError_t fd_launch_inject(float* u, int64_t ncell, const int64_t* src_idx, const float* w, int64_t w_stride,
                            int nshots, hipStream_t st);
hipError_t fd_launch_coef(const float* vel, float* coef, int64_t n, float dt2, float damp, hipStream_t st);
hipError_t fd_launch_vmax(const float* vel, int64_t n, unsigned int* out, hipStream_t st);
hipError_t fd_launch_sponge(float* d, int nxh, int nyh, int nzh, int hw, int sw, float sg, hipStream_t st);

}  // namespace stp
