// 3-D convolution building blocks of the learned preconditioner network
// (BF16 operands on MFMA, FP32 accumulation).  Internal interface.
//
// Activations are channels-last: element (voxel p, channel c) at p*C + c with
// p = (x*n1 + y)*n2 + z.  Convolutions are 3x3x3 with zero padding; tap
// t = (dx+1)*9 + (dy+1)*3 + (dz+1) for offsets dx, dy, dz in {-1, 0, 1}.
// Weights are [cout][27][cin] (the GEMM K index is t*cin + ci).  cin and cout
// must be multiples of 16 (cout at most 96, not 80).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace stp {

typedef uint16_t bf16_t;   // storage type of BF16 values (raw bits)

enum ConvOut {
    CONV_OUT_BF16 = 0,      // BF16 channels-last, row stride out_ld, channel offset out_off
    CONV_OUT_F32 = 1,       // FP32 channels-last, row stride out_ld, channel offset out_off
    CONV_OUT_F32_SOA = 2,   // FP32 [cout][nvox] (channel-major)
};

struct ConvShape {
    int n0, n1, n2;         // voxel grid
    int cin, cout;
};

// out = act(conv(in, w)); relu applies max(0, .) before the store.
hipError_t conv3d_fwd(const ConvShape& s, const bf16_t* in, const bf16_t* w, int relu, int out_kind, void* out,
                      int out_ld, int out_off, hipStream_t st);
// dw[co][t][ci] = sum_p dout[p][co] * in[p + d_t][ci]: FP32 partial sums of
// voxel chunks in `work` (conv3d_bwd_weight_work(s) floats), then summed
// over the chunks in chunk order, so the result does not depend on timing
size_t conv3d_bwd_weight_work(const ConvShape& s);
// an upper bound of conv3d_bwd_weight_work over all shapes
constexpr size_t CONV_BWD_WORK_MAX = (size_t)8192 * 256;
hipError_t conv3d_bwd_weight(const ConvShape& s, const bf16_t* in, const bf16_t* dout, float* dw, float* work,
                             hipStream_t st);
// weights of the data-gradient convolution: wb[ci][t][co] = w[co][26 - t][ci]
hipError_t conv3d_flip_weights(int cin, int cout, const float* w32, bf16_t* wb, hipStream_t st);

// elementwise helpers (channels-last unless stated)
hipError_t f32_to_bf16(int64_t n, const float* a, bf16_t* b, hipStream_t st);
// d_bf16 = bf16(g * (act > 0))
hipError_t relu_mask_bf16(int64_t n, const float* g, const bf16_t* act, bf16_t* d, hipStream_t st);
// 2x2x2 average pooling (n0, n1, n2 even)
hipError_t avgpool2_fwd(int n0, int n1, int n2, int c, const bf16_t* in, bf16_t* out, hipStream_t st);
// din (fine) += dout (coarse) / 8
hipError_t avgpool2_bwd(int n0, int n1, int n2, int c, const float* dout, float* din, hipStream_t st);
// cat[p] = (coarse[parent(p)] (c0 channels), fine[p] (c1 channels)), nearest upsampling by 2
hipError_t upcat_fwd(int n0, int n1, int n2, int c0, int c1, const bf16_t* coarse, const bf16_t* fine, bf16_t* cat,
                     hipStream_t st);
// dcoarse = sum over the 8 children of dcat[:, :c0]; dfine = dcat[:, c0:] (both overwritten)
hipError_t upcat_bwd(int n0, int n1, int n2, int c0, int c1, const float* dcat, float* dcoarse, float* dfine,
                     hipStream_t st);
// Adam step on FP32 master weights; also writes the BF16 copy
hipError_t adam_step(int64_t n, float* w, const float* g, float* m, float* v, float lr, float b1, float b2, float eps,
                     float c1, float c2, bf16_t* wbf, hipStream_t st);

}  // namespace stp
