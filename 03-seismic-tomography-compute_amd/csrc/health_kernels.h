#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace stp {
hipError_t health_count(const double* d, int64_t n, int64_t* out_host, hipStream_t st);
hipError_t health_clamp(double* d, int64_t n, double lo, double hi, hipStream_t st);
hipError_t health_repair(double* d, int64_t count, int64_t nx, int64_t ny, int64_t nz, int hw, double repl,
                         hipStream_t st);
hipError_t field_energy_f32(const float* u, int64_t n, double* out_host, hipStream_t st);
hipError_t field_interior_f64(const float* u, int64_t nxh, int64_t nyh, int64_t nzh, int hw, double* out,
                              hipStream_t st);
}  // namespace stp
