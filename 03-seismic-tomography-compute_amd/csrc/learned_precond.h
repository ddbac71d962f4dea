// Learned preconditioner (3-level 3-D convolutional network) used by CGLS.
// Internal interface.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace stp {
// z = M s for the system prepared in `handle` (an stp_lpc*); kernels return
// immediately when *done != 0.  With s_prev, part_zs and part_fl (all three or
// none) it also writes the FP64 partial sums of z.s and z.(s - s_prev), one
// per block of 256 cells.
// z = M s.  With perm (cell -> position of that cell's value in s, z and
// s_prev) the vectors are stored in that order; M acts on cells.
int lpc_apply(void* handle, const float* s, float* z, const int* done, const float* s_prev, double* part_zs,
              double* part_fl, hipStream_t st, const int* perm = nullptr);
// number of cells of the prepared system, -1 when none is prepared
int64_t lpc_cols(const void* handle);
}  // namespace stp
