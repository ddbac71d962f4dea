// Solver layout of a tomography matrix (solver_layout.hip), internal.
#pragma once
#include <hip/hip_runtime.h>

#include "csr_handle.h"

namespace stp {
int build_solver_layout(stp_csr* h, hipStream_t st);   // no-op without a grid or when built
void free_solver_layout(stp_csr* h);
hipError_t gather_f64(int n, const double* src, const int* idx, double* dst, hipStream_t st);   // dst[i] = src[idx[i]]
}  // namespace stp
