// Solver layout of a tomography matrix (solver_layout.hip), internal.
#pragma once
#include <hip/hip_runtime.h>

#include "csr_handle.h"

namespace stp {
int build_solver_layout(stp_csr* h, hipStream_t st);   // no-op without a grid or when built
void free_solver_layout(stp_csr* h);
// dst[idx[i]] = src[i] (scatter) or dst[i] = src[idx[i]] (gather), i < n
template <typename T>
hipError_t permute(int n, const T* src, const int* idx, T* dst, bool scatter, hipStream_t st);
}  // namespace stp
