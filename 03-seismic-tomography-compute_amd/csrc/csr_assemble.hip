// Canonical CSR assembly and the FP64 CSR utilities behind the CSRMatrix
// drop-in (src/cpp/sparse/csr_matrix.cpp:25-193).
//
// Assembly = CSRMatrix::build_from_triplets semantics without its padding
// pass: triplets with row or column out of range are dropped, each row is
// ordered by column, equal (row, col) entries are summed in input order.  On
// the GPU this is one stable radix sort of (row*cols + col) keys carrying the
// input position, a head flag per distinct key, and an in-order segment sum.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>

#include "capi_common.h"

namespace stp {
namespace {

__global__ void asm_keys_kernel(int64_t n, int64_t rows, int64_t cols, const int64_t* r, const int64_t* c,
                                uint64_t* keys, int64_t* pos, unsigned long long* valid) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool ok = r[i] >= 0 && r[i] < rows && c[i] >= 0 && c[i] < cols;
    keys[i] = ok ? (uint64_t)r[i] * (uint64_t)cols + (uint64_t)c[i] : ~0ull;
    pos[i] = i;
    if (ok) atomicAdd(valid, 1ull);
}

__global__ void asm_heads_kernel(int64_t m, const uint64_t* keys, int64_t* head) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    head[k] = (k == 0 || keys[k] != keys[k - 1]) ? 1 : 0;
}

__global__ void asm_merge_kernel(int64_t m, int64_t cols, const uint64_t* keys, const int64_t* pos, const double* v,
                                 const int64_t* head, const int64_t* slot, int64_t* col_out, double* val_out,
                                 unsigned long long* row_cnt) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m || !head[k]) return;
    const uint64_t key = keys[k];
    double s = v[pos[k]];
    for (int64_t j = k + 1; j < m && keys[j] == key; ++j) s += v[pos[j]];
    const int64_t o = slot[k];
    col_out[o] = (int64_t)(key % (uint64_t)cols);
    val_out[o] = s;
    atomicAdd(&row_cnt[key / (uint64_t)cols], 1ull);
}

__global__ void u64_to_i64_kernel(int64_t n, const unsigned long long* a, int64_t* b) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (int64_t)a[i];
}

__global__ void scale_kernel(int64_t n, double* v, double a) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        v[i] *= a;
}

__global__ void add_diag_kernel(int64_t rows, const int64_t* rp, const int64_t* ci, double* v, double a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    for (int64_t p = rp[i]; p < rp[i + 1]; ++p)
        if (ci[p] == i) { v[p] += a; break; }
}

__global__ void sumsq_kernel(int64_t n, const double* a, const double* b, double* part) {
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        s += a[i] * b[i];
    __shared__ double red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void to_f32_kernel(int64_t rows, int64_t nnz, const int64_t* rp, const int64_t* ci, const double* v,
                              int* rp32, int* ci32, float* v32) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= rows) rp32[i] = (int)rp[i];
    if (i < nnz) { ci32[i] = (int)ci[i]; v32[i] = (float)v[i]; }
}

// 7-point FD matrix triplets (build_fd_stencil_7point, csr_matrix.cpp:74-116);
// missing neighbours get row -1 and are dropped by the assembly
__global__ void fd7_triplets_kernel(int64_t nx, int64_t ny, int64_t nz, const double* vel, double idx2, double idy2,
                                    double idz2, double dt2, int64_t* r, int64_t* c, double* v) {
    const int64_t n = nx * ny * nz;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t ix = i / (ny * nz), iy = (i / nz) % ny, iz = i % nz;
    const double vv = vel[i];
    const double v2dt2 = vv * vv * dt2;
    const int64_t b = 7 * i;
    r[b] = i; c[b] = i; v[b] = -2.0 * v2dt2 * (idx2 + idy2 + idz2) + 2.0;
    r[b + 1] = ix > 0 ? i : -1;        c[b + 1] = i - ny * nz; v[b + 1] = v2dt2 * idx2;
    r[b + 2] = ix < nx - 1 ? i : -1;   c[b + 2] = i + ny * nz; v[b + 2] = v2dt2 * idx2;
    r[b + 3] = iy > 0 ? i : -1;        c[b + 3] = i - nz;      v[b + 3] = v2dt2 * idy2;
    r[b + 4] = iy < ny - 1 ? i : -1;   c[b + 4] = i + nz;      v[b + 4] = v2dt2 * idy2;
    r[b + 5] = iz > 0 ? i : -1;        c[b + 5] = i - 1;       v[b + 5] = v2dt2 * idz2;
    r[b + 6] = iz < nz - 1 ? i : -1;   c[b + 6] = i + 1;       v[b + 6] = v2dt2 * idz2;
}

unsigned g256(int64_t n) {
    int64_t b = (n + 255) / 256;
    return (unsigned)(b < 1 ? 1 : b);
}
unsigned g256c(int64_t n) {
    int64_t b = (n + 255) / 256;
    if (b > 4096) b = 4096;
    return (unsigned)(b < 1 ? 1 : b);
}

}  // namespace
}  // namespace stp

using stp::fail;
using stp::S;

extern "C" {

int stp_csr_assemble(int64_t rows, int64_t cols, int64_t n, const int64_t* r, const int64_t* c, const double* v,
                     int64_t* row_ptr, int64_t* col_out, double* val_out, int64_t* nnz_host, void* stream) {
    if (rows < 0 || cols < 0 || n < 0 || !row_ptr || !nnz_host || (n > 0 && (!r || !c || !v || !col_out || !val_out)))
        return fail(STP_ERR_INVALID, "stp_csr_assemble: bad argument");
    if (rows > 0x7fffffff || cols > 0x7fffffff || n > 0x7fffffff)
        return fail(STP_ERR_INVALID, "stp_csr_assemble: sizes must be below 2^31");
    hipStream_t st = S(stream);
    *nnz_host = 0;
    uint64_t *k0 = nullptr, *k1 = nullptr;
    int64_t *p0 = nullptr, *p1 = nullptr, *head = nullptr, *slot = nullptr, *cnt64 = nullptr;
    unsigned long long *valid = nullptr, *rcnt = nullptr;
    void* tmp = nullptr;
    hipError_t e = hipSuccess;
    int64_t m = 0, nnz = 0;
    do {
        if ((e = hipMalloc(&rcnt, sizeof(unsigned long long) * (rows + 1))) != hipSuccess) break;
        if ((e = hipMemsetAsync(rcnt, 0, sizeof(unsigned long long) * (rows + 1), st)) != hipSuccess) break;
        if ((e = hipMalloc(&cnt64, sizeof(int64_t) * (rows + 1))) != hipSuccess) break;
        if (n > 0) {
            if ((e = hipMalloc(&k0, 8 * n)) != hipSuccess || (e = hipMalloc(&k1, 8 * n)) != hipSuccess ||
                (e = hipMalloc(&p0, 8 * n)) != hipSuccess || (e = hipMalloc(&p1, 8 * n)) != hipSuccess ||
                (e = hipMalloc(&head, 8 * n)) != hipSuccess || (e = hipMalloc(&slot, 8 * n)) != hipSuccess ||
                (e = hipMalloc(&valid, 8)) != hipSuccess)
                break;
            if ((e = hipMemsetAsync(valid, 0, 8, st)) != hipSuccess) break;
            hipLaunchKernelGGL(stp::asm_keys_kernel, dim3(stp::g256(n)), dim3(256), 0, st, n, rows, cols, r, c, k0,
                               p0, valid);
            if ((e = hipGetLastError()) != hipSuccess) break;
            size_t tb = 0;
            if ((e = hipcub::DeviceRadixSort::SortPairs(nullptr, tb, k0, k1, p0, p1, (int)n, 0, 64, st)) != hipSuccess)
                break;
            if ((e = hipMalloc(&tmp, tb ? tb : 1)) != hipSuccess) break;
            if ((e = hipcub::DeviceRadixSort::SortPairs(tmp, tb, k0, k1, p0, p1, (int)n, 0, 64, st)) != hipSuccess)
                break;
            unsigned long long hv = 0;
            if ((e = hipMemcpyAsync(&hv, valid, 8, hipMemcpyDeviceToHost, st)) != hipSuccess) break;
            if ((e = hipStreamSynchronize(st)) != hipSuccess) break;
            m = (int64_t)hv;
            if (m > 0) {
                hipLaunchKernelGGL(stp::asm_heads_kernel, dim3(stp::g256(m)), dim3(256), 0, st, m, k1, head);
                (void)hipFree(tmp);
                tmp = nullptr;
                tb = 0;
                if ((e = hipcub::DeviceScan::ExclusiveSum(nullptr, tb, head, slot, (int)m, st)) != hipSuccess) break;
                if ((e = hipMalloc(&tmp, tb ? tb : 1)) != hipSuccess) break;
                if ((e = hipcub::DeviceScan::ExclusiveSum(tmp, tb, head, slot, (int)m, st)) != hipSuccess) break;
                int64_t last_slot = 0, last_head = 0;
                if ((e = hipMemcpyAsync(&last_slot, slot + m - 1, 8, hipMemcpyDeviceToHost, st)) != hipSuccess) break;
                if ((e = hipMemcpyAsync(&last_head, head + m - 1, 8, hipMemcpyDeviceToHost, st)) != hipSuccess) break;
                if ((e = hipStreamSynchronize(st)) != hipSuccess) break;
                nnz = last_slot + last_head;
                hipLaunchKernelGGL(stp::asm_merge_kernel, dim3(stp::g256(m)), dim3(256), 0, st, m, cols, k1, p1, v,
                                   head, slot, col_out, val_out, rcnt);
                if ((e = hipGetLastError()) != hipSuccess) break;
            }
        }
        hipLaunchKernelGGL(stp::u64_to_i64_kernel, dim3(stp::g256(rows + 1)), dim3(256), 0, st, rows + 1, rcnt, cnt64);
        if ((e = hipGetLastError()) != hipSuccess) break;
        if (tmp) { (void)hipFree(tmp); tmp = nullptr; }
        size_t tb = 0;
        if ((e = hipcub::DeviceScan::ExclusiveSum(nullptr, tb, cnt64, row_ptr, (int)(rows + 1), st)) != hipSuccess)
            break;
        if ((e = hipMalloc(&tmp, tb ? tb : 1)) != hipSuccess) break;
        if ((e = hipcub::DeviceScan::ExclusiveSum(tmp, tb, cnt64, row_ptr, (int)(rows + 1), st)) != hipSuccess) break;
        e = hipStreamSynchronize(st);
    } while (0);
    (void)hipFree(tmp); (void)hipFree(k0); (void)hipFree(k1); (void)hipFree(p0); (void)hipFree(p1); (void)hipFree(head); (void)hipFree(slot);
    (void)hipFree(valid); (void)hipFree(rcnt); (void)hipFree(cnt64);
    if (e != hipSuccess) return stp::hip_fail(e, "stp_csr_assemble");
    *nnz_host = nnz;
    return STP_OK;
}

int stp_csr64_fd7_triplets(int64_t nx, int64_t ny, int64_t nz, const double* vel, double dx, double dy, double dz,
                           double dt, int64_t* r, int64_t* c, double* v, void* stream) {
    if (nx < 1 || ny < 1 || nz < 1 || !vel || !r || !c || !v) return fail(STP_ERR_INVALID, "stp_csr64_fd7_triplets: bad argument");
    const int64_t n = nx * ny * nz;
    hipLaunchKernelGGL(stp::fd7_triplets_kernel, dim3(stp::g256(n)), dim3(256), 0, S(stream), nx, ny, nz, vel,
                       1.0 / (dx * dx), 1.0 / (dy * dy), 1.0 / (dz * dz), dt * dt, r, c, v);
    STP_HIP(hipGetLastError());
    return STP_OK;
}

int stp_csr64_scale(int64_t nnz, double* val, double alpha, void* stream) {
    if (nnz < 0 || (nnz && !val)) return fail(STP_ERR_INVALID, "stp_csr64_scale: bad argument");
    if (nnz) hipLaunchKernelGGL(stp::scale_kernel, dim3(stp::g256c(nnz)), dim3(256), 0, S(stream), nnz, val, alpha);
    STP_HIP(hipGetLastError());
    return STP_OK;
}

int stp_csr64_add_diagonal(int64_t rows, const int64_t* rp, const int64_t* ci, double* val, double alpha,
                           void* stream) {
    if (rows < 0 || (rows && (!rp || !ci || !val))) return fail(STP_ERR_INVALID, "stp_csr64_add_diagonal: bad argument");
    if (rows)
        hipLaunchKernelGGL(stp::add_diag_kernel, dim3(stp::g256(rows)), dim3(256), 0, S(stream), rows, rp, ci, val,
                           alpha);
    STP_HIP(hipGetLastError());
    return STP_OK;
}

/* sum_i a_i * b_i in FP64 (dot products, squared norms) */
int stp_dot_f64(int64_t n, const double* a, const double* b, double* out_host, void* stream) {
    if (n < 0 || !out_host || (n && (!a || !b))) return fail(STP_ERR_INVALID, "stp_dot_f64: bad argument");
    *out_host = 0.0;
    if (!n) return STP_OK;
    hipStream_t st = S(stream);
    const unsigned nb = stp::g256c(n);
    double* part = nullptr;
    STP_HIP(hipMalloc(&part, sizeof(double) * nb));
    hipLaunchKernelGGL(stp::sumsq_kernel, dim3(nb), dim3(256), 0, st, n, a, b, part);
    hipError_t e = hipGetLastError();
    double* hp = new double[nb];
    if (e == hipSuccess) e = hipMemcpyAsync(hp, part, sizeof(double) * nb, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(part);
    double s = 0.0;
    for (unsigned i = 0; i < nb; ++i) s += hp[i];
    delete[] hp;
    if (e != hipSuccess) return stp::hip_fail(e, "stp_dot_f64");
    *out_host = s;
    return STP_OK;
}

/* FP64/int64 CSR -> device FP32/int32 arrays (nnz < 2^31) */
int stp_csr64_to_f32(int64_t rows, int64_t nnz, const int64_t* rp, const int64_t* ci, const double* val,
                     int32_t* rp32, int32_t* ci32, float* val32, void* stream) {
    if (rows < 0 || nnz < 0 || nnz > 0x7fffffff || !rp || !rp32 || (nnz && (!ci || !val || !ci32 || !val32)))
        return fail(STP_ERR_INVALID, "stp_csr64_to_f32: bad argument");
    const int64_t n = (rows + 1 > nnz ? rows + 1 : nnz);
    hipLaunchKernelGGL(stp::to_f32_kernel, dim3(stp::g256(n)), dim3(256), 0, S(stream), rows, nnz, rp, ci, val, rp32,
                       ci32, val32);
    STP_HIP(hipGetLastError());
    return STP_OK;
}

}  // extern "C"
