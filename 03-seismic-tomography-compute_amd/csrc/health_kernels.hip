// Field-health kernels: GPU versions of FaultTolerance::count_nan_inf,
// clamp_field and repair_field_nan (src/cpp/core/fault_tolerance.cpp:72-157).
// These are rare-path utilities exposed through check_field_health /
// clamp_field_safe / repair_field_nan (src/cpp/pybind/bindings.cpp:275-301);
// the propagator itself never produces non-finite values (its guards replace
// them), so they are not on the per-step path.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>

#include "health_kernels.h"

namespace stp {

__global__ void count_nonfinite_kernel(const double* d, int64_t n, unsigned long long* out) {
    unsigned long long c = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        c += !__builtin_isfinite(d[i]);
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(out, c);
}

__global__ void clamp_kernel(double* d, int64_t n, double lo, double hi) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double v = d[i];
        d[i] = !__builtin_isfinite(v) ? 0.0 : fmax(lo, fmin(v, hi));
    }
}

// interior non-finite cells -> unordered list
__global__ void collect_bad_kernel(const double* d, int64_t nx, int64_t ny, int64_t nz, int hw,
                                   unsigned long long* cnt, int64_t* list) {
    const int64_t nix = nx - 2 * hw, niy = ny - 2 * hw, niz = nz - 2 * hw;
    if (nix <= 0 || niy <= 0 || niz <= 0) return;
    const int64_t n = nix * niy * niz;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t ix = i / (niy * niz) + hw, iy = (i / niz) % niy + hw, iz = i % niz + hw;
        const int64_t c = (ix * ny + iy) * nz + iz;
        if (!__builtin_isfinite(d[c])) list[atomicAdd(cnt, 1ull)] = c;
    }
}

// sequential neighbour-mean repair in increasing index order: a repaired
// cell is visible to the cells after it, exactly as in the reference loop.
__global__ void repair_sequential_kernel(double* d, const int64_t* list, const unsigned long long* cnt, int64_t nx,
                                         int64_t ny, int64_t nz, int hw, double repl) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const int64_t nynz = ny * nz;
    const int64_t m = (int64_t)*cnt;
    for (int64_t k = 0; k < m; ++k) {
        const int64_t i = list[k];
        const int64_t ix = i / nynz, iy = (i / nz) % ny, iz = i % nz;
        double s = 0.0;
        int c = 0;
        if (ix > hw && __builtin_isfinite(d[i - nynz])) { s += d[i - nynz]; ++c; }
        if (ix < nx - hw - 1 && __builtin_isfinite(d[i + nynz])) { s += d[i + nynz]; ++c; }
        if (iy > hw && __builtin_isfinite(d[i - nz])) { s += d[i - nz]; ++c; }
        if (iy < ny - hw - 1 && __builtin_isfinite(d[i + nz])) { s += d[i + nz]; ++c; }
        if (iz > hw && __builtin_isfinite(d[i - 1])) { s += d[i - 1]; ++c; }
        if (iz < nz - hw - 1 && __builtin_isfinite(d[i + 1])) { s += d[i + 1]; ++c; }
        d[i] = c > 0 ? s / c : repl;
    }
}

__global__ void replace_nonfinite_kernel(double* d, int64_t n, double repl) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        if (!__builtin_isfinite(d[i])) d[i] = repl;
}

// WaveEquationSolver::compute_energy (wave_equation.cpp:482-490): sum of u^2
// over the finite values of an FP32 field, FP64 partial sums per block (each
// product of two FP32 values is exact in FP64)
__global__ void energy_f32_kernel(const float* u, int64_t n, double* part) {
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double v = (double)u[i];
        if (__builtin_isfinite(v)) s += v * v;
    }
    __shared__ double red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// WaveEquationSolver::save_snapshot (wave_equation.cpp:463-480): the interior
// of a with-halo FP32 field as FP64, x-major, non-finite values as 0
__global__ void interior_f64_kernel(const float* u, int64_t nxh, int64_t nyh, int64_t nzh, int hw, double* out) {
    const int64_t nx = nxh - 2 * hw, ny = nyh - 2 * hw, nz = nzh - 2 * hw;
    const int64_t n = nx * ny * nz;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t ix = i / (ny * nz), iy = (i / nz) % ny, iz = i % nz;
        const double v = (double)u[((ix + hw) * nyh + (iy + hw)) * nzh + (iz + hw)];
        out[i] = __builtin_isfinite(v) ? v : 0.0;
    }
}

static unsigned grid_for(int64_t n) {
    int64_t b = (n + 255) / 256;
    if (b > 8192) b = 8192;
    if (b < 1) b = 1;
    return (unsigned)b;
}

hipError_t health_count(const double* d, int64_t n, int64_t* out_host, hipStream_t st) {
    unsigned long long* dc = nullptr;
    hipError_t e = hipMalloc(&dc, sizeof(*dc));
    if (e != hipSuccess) return e;
    unsigned long long h = 0;
    if ((e = hipMemsetAsync(dc, 0, sizeof(*dc), st)) == hipSuccess) {
        hipLaunchKernelGGL(count_nonfinite_kernel, dim3(grid_for(n)), dim3(256), 0, st, d, n, dc);
        if ((e = hipGetLastError()) == hipSuccess &&
            (e = hipMemcpyAsync(&h, dc, sizeof(h), hipMemcpyDeviceToHost, st)) == hipSuccess)
            e = hipStreamSynchronize(st);
    }
    (void)hipFree(dc);
    *out_host = (int64_t)h;
    return e;
}

hipError_t health_clamp(double* d, int64_t n, double lo, double hi, hipStream_t st) {
    hipLaunchKernelGGL(clamp_kernel, dim3(grid_for(n)), dim3(256), 0, st, d, n, lo, hi);
    return hipGetLastError();
}

hipError_t health_repair(double* d, int64_t count, int64_t nx, int64_t ny, int64_t nz, int hw, double repl,
                         hipStream_t st) {
    const int64_t nint = (nx - 2 * hw) * (ny - 2 * hw) * (nz - 2 * hw);
    hipError_t e = hipSuccess;
    if (nx > 2 * hw && ny > 2 * hw && nz > 2 * hw && nx * ny * nz <= count) {
        int64_t bad = 0;
        if ((e = health_count(d, count, &bad, st)) != hipSuccess) return e;
        if (bad > 0) {
            const int64_t cap = bad < nint ? bad : nint;
            unsigned long long* dc = nullptr;
            int64_t *l0 = nullptr, *l1 = nullptr;
            void* tmp = nullptr;
            size_t tb = 0;
            if ((e = hipMalloc(&dc, sizeof(*dc))) == hipSuccess && (e = hipMalloc(&l0, cap * 8)) == hipSuccess &&
                (e = hipMalloc(&l1, cap * 8)) == hipSuccess &&
                (e = hipMemsetAsync(dc, 0, sizeof(*dc), st)) == hipSuccess) {
                hipLaunchKernelGGL(collect_bad_kernel, dim3(grid_for(nint)), dim3(256), 0, st, d, nx, ny, nz, hw, dc,
                                   l0);
                unsigned long long found = 0;
                if ((e = hipGetLastError()) == hipSuccess &&
                    (e = hipMemcpyAsync(&found, dc, sizeof(found), hipMemcpyDeviceToHost, st)) == hipSuccess &&
                    (e = hipStreamSynchronize(st)) == hipSuccess && found > 0) {
                    e = hipcub::DeviceRadixSort::SortKeys(nullptr, tb, l0, l1, (int)found, 0, 64, st);
                    if (e == hipSuccess && (e = hipMalloc(&tmp, tb ? tb : 1)) == hipSuccess)
                        e = hipcub::DeviceRadixSort::SortKeys(tmp, tb, l0, l1, (int)found, 0, 64, st);
                    if (e == hipSuccess) {
                        hipLaunchKernelGGL(repair_sequential_kernel, dim3(1), dim3(1), 0, st, d, l1, dc, nx, ny, nz,
                                           hw, repl);
                        e = hipGetLastError();
                    }
                    if (e == hipSuccess) e = hipStreamSynchronize(st);
                }
            }
            if (tmp) (void)hipFree(tmp);
            if (l1) (void)hipFree(l1);
            if (l0) (void)hipFree(l0);
            if (dc) (void)hipFree(dc);
            if (e != hipSuccess) return e;
        }
    }
    hipLaunchKernelGGL(replace_nonfinite_kernel, dim3(grid_for(count)), dim3(256), 0, st, d, count, repl);
    return hipGetLastError();
}

hipError_t field_energy_f32(const float* u, int64_t n, double* out_host, hipStream_t st) {
    *out_host = 0.0;
    if (n <= 0) return hipSuccess;
    const unsigned nb = grid_for(n) > 1024 ? 1024 : grid_for(n);
    double* part = nullptr;
    hipError_t e = hipMalloc(&part, sizeof(double) * nb);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(energy_f32_kernel, dim3(nb), dim3(256), 0, st, u, n, part);
    e = hipGetLastError();
    double hp[1024];
    if (e == hipSuccess) e = hipMemcpyAsync(hp, part, sizeof(double) * nb, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(part);
    if (e != hipSuccess) return e;
    double sum = 0.0;
    for (unsigned k = 0; k < nb; ++k) sum += hp[k];
    *out_host = sum;
    return hipSuccess;
}

hipError_t field_interior_f64(const float* u, int64_t nxh, int64_t nyh, int64_t nzh, int hw, double* out,
                              hipStream_t st) {
    const int64_t n = (nxh - 2 * hw) * (nyh - 2 * hw) * (nzh - 2 * hw);
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(interior_f64_kernel, dim3(grid_for(n)), dim3(256), 0, st, u, nxh, nyh, nzh, hw, out);
    return hipGetLastError();
}

}  // namespace stp
