// Learned preconditioner (3-level 3-D convolutional network) used by the
// flexible CGLS.  Internal interface.
#pragma once
#include <hip/hip_runtime.h>

namespace stp {
// z = M(s) for the network held by `handle` (an stp_lpc*); kernels return
// immediately when *done != 0.
int lpc_apply(void* handle, const float* s, float* z, const int* done, hipStream_t st);
}  // namespace stp
