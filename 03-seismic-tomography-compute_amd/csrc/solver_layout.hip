// Solver layout of a tomography matrix: a copy of A and A^T renumbered for
// gather locality, used by CGLS when the matrix's columns are the cells of an
// nx*ny*nz grid (x-major, z fastest).
//
//   cells  are renumbered in Morton (bit-interleaved x, y, z) order, so cells
//          that are close in space share cache lines of p, s, z and x;
//   rows   (rays) are sorted by the 6-D Morton code of the two cells that
//          hold the row's smallest and largest column (the ends of a straight
//          ray), so rays that run close together are adjacent and the
//          A^T r gather of a cell reads neighbouring entries of r.
//
// Each row keeps its entries in their original order (only the column ids
// are renumbered), so A p in the layout is A p of the original matrix, bit for
// bit, permuted; A^T is rebuilt from the renumbered A.  The layout costs one
// more copy of A and A^T and is built once per matrix.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

#include "capi_common.h"
#include "csr_handle.h"
#include "seismic_hip.h"
#include "solver_layout.h"
#include "sparse_kernels.h"

namespace stp {
namespace {

unsigned blocks(int64_t n) {
    const int64_t b = (n + 255) / 256;
    return (unsigned)(b < 1 ? 1 : b);
}

int bits_for(int64_t n) {   // bits of n - 1 (0 for n <= 1)
    int b = 0;
    while (b < 62 && ((int64_t)1 << b) < n) ++b;
    return b;
}

// spread the low `nb` bits of v so that bit k lands at bit k * stride
__device__ __forceinline__ uint64_t spread(uint64_t v, int nb, int stride) {
    uint64_t o = 0;
    for (int k = 0; k < nb; ++k) o |= ((v >> k) & 1ull) << (k * stride);
    return o;
}

__global__ void cell_code_kernel(int64_t nx, int64_t ny, int64_t nz, int nb, int shift, uint64_t* key, int* idx) {
    const int64_t n = nx * ny * nz;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t z = (uint64_t)(i % nz) >> shift, y = (uint64_t)((i / nz) % ny) >> shift,
                   x = (uint64_t)(i / (ny * nz)) >> shift;
    key[i] = (spread(x, nb, 3) << 2) | (spread(y, nb, 3) << 1) | spread(z, nb, 3);
    idx[i] = (int)i;
}

__global__ void rank_kernel(int n, const int* sorted_idx, int* perm) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) perm[sorted_idx[k]] = k;
}

// row key: 6-D Morton code of the (x, y, z) of the row's smallest and largest
// column, each coordinate scaled to 10 bits; empty rows get key 0
__global__ void row_key_kernel(int rows, const int* __restrict__ rp, const int* __restrict__ ci, int64_t nx,
                               int64_t ny, int64_t nz, uint64_t* key, int* idx) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const int b = rp[r], e = rp[r + 1];
    uint64_t k = 0;
    if (e > b) {
        int lo = ci[b], hi = lo;
        for (int j = b + 1; j < e; ++j) {
            const int c = ci[j];
            lo = c < lo ? c : lo;
            hi = c > hi ? c : hi;
        }
        const int64_t cs[2] = {lo, hi};
        uint64_t q[6];
        for (int t = 0; t < 2; ++t) {
            const int64_t c = cs[t];
            q[3 * t + 0] = (uint64_t)(((c / (ny * nz)) << 10) / nx);
            q[3 * t + 1] = (uint64_t)((((c / nz) % ny) << 10) / ny);
            q[3 * t + 2] = (uint64_t)(((c % nz) << 10) / nz);
        }
        for (int d = 0; d < 6; ++d) k |= spread(q[d], 10, 6) << (5 - d);
    }
    key[r] = k;
    idx[r] = r;
}

__global__ void row_len_kernel(int rows, const int* __restrict__ rsrc, const int* __restrict__ rp, int* len) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < rows) len[k] = rp[rsrc[k] + 1] - rp[rsrc[k]];
    if (k == rows) len[k] = 0;
}

// one wave per layout row: copy the source row with renumbered columns
__global__ void fill_kernel(int rows, const int* __restrict__ rsrc, const int* __restrict__ rp,
                            const int* __restrict__ ci, const float* __restrict__ val, const int* __restrict__ cperm,
                            const int* __restrict__ lrp, int* __restrict__ lci, float* __restrict__ lval) {
    const int k = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    if (k >= rows) return;
    const int o = rsrc[k], b = rp[o], e = rp[o + 1], d = lrp[k] - b;
    for (int j = b + lane; j < e; j += 64) {
        lci[j + d] = cperm[ci[j]];
        lval[j + d] = val[j];
    }
}

// staged A^T: key (tile, column) of every entry, one wave per row
__global__ void stage_key_kernel(int rows, const int* __restrict__ rp, const int* __restrict__ ci, int rt,
                                 uint64_t* __restrict__ key, int* __restrict__ idx) {
    const int k = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    if (k >= rows) return;
    const uint64_t t = (uint64_t)(k / rt) << 32;
    for (int j = rp[k] + lane; j < rp[k + 1]; j += 64) {
        key[j] = t | (uint32_t)ci[j];
        idx[j] = j;
    }
}

// after the sort: flag the first entry of each distinct key (flag[nnz] = 0)
__global__ void stage_flag_kernel(int nnz, const uint64_t* __restrict__ key, int* __restrict__ flag) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j > nnz) return;
    flag[j] = (j < nnz && (j == 0 || key[j] != key[j - 1])) ? 1 : 0;
}

// uidx = exclusive scan of flag.  The sorted entries of tile t start at
// rp[t * rt] (the entries of the tiles before it), so its list starts at
// tu[t] = uidx[rp[min(t * rt, rows)]]; tu[nt] is the total.
__global__ void stage_tiles_kernel(int nt, int rt, int rows, const int* __restrict__ rp, const int* __restrict__ uidx,
                                   int* __restrict__ tu) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > nt) return;
    const int64_t r = (int64_t)t * rt < rows ? (int64_t)t * rt : rows;
    tu[t] = uidx[rp[r]];
}

// the entry's distinct key is number uidx + flag - 1, its tile's list starts at tu
__global__ void stage_fill_kernel(int nnz, const uint64_t* __restrict__ key, const int* __restrict__ flag,
                                  const int* __restrict__ uidx, const int* __restrict__ idx,
                                  const int* __restrict__ tu, int* __restrict__ uniq, uint16_t* __restrict__ l16) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nnz) return;
    const int t = (int)(key[j] >> 32);
    const int u = uidx[j] + flag[j] - 1;
    if (flag[j]) uniq[u] = (int)(uint32_t)key[j];
    l16[idx[j]] = (uint16_t)(u - tu[t]);
}

template <typename K>
hipError_t sort_pairs(K* k0, K* k1, int* v0, int* v1, int n, int end_bit, hipStream_t st) {
    size_t tb = 0;
    void* tmp = nullptr;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, tb, k0, k1, v0, v1, n, 0, end_bit, st);
    if (e == hipSuccess) e = hipMalloc(&tmp, tb ? tb : 1);
    if (e == hipSuccess) e = hipcub::DeviceRadixSort::SortPairs(tmp, tb, k0, k1, v0, v1, n, 0, end_bit, st);
    if (tmp) {
        const hipError_t e2 = hipStreamSynchronize(st);
        (void)hipFree(tmp);
        if (e == hipSuccess) e = e2;
    }
    return e;
}

// Staged copy of the layout A^T (n rows, entries reading a vector of length
// g): tiles of 64 rows, or 32 when a 64-row tile would read more than
// STAGE_CAP distinct entries of that vector.  Taller tiles reuse more (at
// 200k rays / 64^3 each staged ray serves 4.0 entries at 64 rows, 4.8 at
// 128) but give half as many workgroups to balance: CGLS took 34.9 ms at 64
// rows, 36.7 at 128, 42.9 at 256.  None fitting leaves st_rt = 0 (the CGLS
// iteration then gathers from global memory).
hipError_t build_staged_transpose(stp_csr* h, hipStream_t st) {
    const int n = h->cols;
    const int nnz = (int)h->nnz;
    if (n == 0 || nnz == 0) return hipSuccess;
    uint64_t *k0 = nullptr, *k1 = nullptr;
    int *i0 = nullptr, *i1 = nullptr, *flag = nullptr, *uidx = nullptr, *tu = nullptr;
    hipError_t e = hipSuccess;
    do {
        if ((e = hipMalloc(&k0, sizeof(uint64_t) * nnz)) != hipSuccess) break;
        if ((e = hipMalloc(&k1, sizeof(uint64_t) * nnz)) != hipSuccess) break;
        if ((e = hipMalloc(&i0, sizeof(int) * nnz)) != hipSuccess) break;
        if ((e = hipMalloc(&i1, sizeof(int) * nnz)) != hipSuccess) break;
        if ((e = hipMalloc(&flag, sizeof(int) * ((size_t)nnz + 1))) != hipSuccess) break;
        if ((e = hipMalloc(&uidx, sizeof(int) * ((size_t)nnz + 1))) != hipSuccess) break;
        for (int rt = 64; rt >= 32; rt /= 2) {
            const int nt = (n + rt - 1) / rt;
            (void)hipFree(tu);
            tu = nullptr;
            if ((e = hipMalloc(&tu, sizeof(int) * ((size_t)nt + 1))) != hipSuccess) break;
            hipLaunchKernelGGL(stage_key_kernel, dim3(blocks((int64_t)n * 64)), dim3(256), 0, st, n, h->ltrp, h->ltci,
                               rt, k0, i0);
            if ((e = hipGetLastError()) != hipSuccess) break;
            if ((e = sort_pairs(k0, k1, i0, i1, nnz, 32 + bits_for((int64_t)nt), st)) != hipSuccess) break;
            hipLaunchKernelGGL(stage_flag_kernel, dim3(blocks((int64_t)nnz + 1)), dim3(256), 0, st, nnz, k1, flag);
            if ((e = hipGetLastError()) != hipSuccess) break;
            if ((e = exclusive_scan_i32(flag, uidx, nnz + 1, st)) != hipSuccess) break;
            hipLaunchKernelGGL(stage_tiles_kernel, dim3(blocks((int64_t)nt + 1)), dim3(256), 0, st, nt, rt, n, h->ltrp,
                               uidx, tu);
            if ((e = hipGetLastError()) != hipSuccess) break;
            std::vector<int> htu((size_t)nt + 1);
            if ((e = hipMemcpyAsync(htu.data(), tu, sizeof(int) * ((size_t)nt + 1), hipMemcpyDeviceToHost, st)) !=
                hipSuccess)
                break;
            if ((e = hipStreamSynchronize(st)) != hipSuccess) break;
            int maxu = 0;
            for (int t = 0; t < nt; ++t) maxu = std::max(maxu, htu[(size_t)t + 1] - htu[(size_t)t]);
            if (maxu > STAGE_CAP) continue;
            if ((e = hipMalloc(&h->st_uniq, sizeof(int) * (size_t)std::max(htu[(size_t)nt], 1))) != hipSuccess) break;
            if ((e = hipMalloc(&h->st_l16, sizeof(uint16_t) * (size_t)nnz)) != hipSuccess) break;
            hipLaunchKernelGGL(stage_fill_kernel, dim3(blocks(nnz)), dim3(256), 0, st, nnz, k1, flag, uidx, i1, tu,
                               h->st_uniq, h->st_l16);
            if ((e = hipGetLastError()) != hipSuccess) break;
            h->st_tu = tu;
            tu = nullptr;
            h->st_rt = rt;
            h->st_ntiles = nt;
            h->st_maxu = maxu;
            h->st_nlist = htu[(size_t)nt];
            break;
        }
        if (e == hipSuccess) e = hipStreamSynchronize(st);
    } while (0);
    (void)hipFree(k0); (void)hipFree(k1); (void)hipFree(i0); (void)hipFree(i1); (void)hipFree(flag); (void)hipFree(uidx); (void)hipFree(tu);
    return e;
}

}  // namespace

void free_solver_layout(stp_csr* h) {
    void* ps[] = {h->cperm, h->rsrc, h->lrp, h->lci, h->lval, h->ltrp, h->ltci, h->ltval, h->lx, h->lb,
                  h->st_tu, h->st_uniq, h->st_l16};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    h->cperm = h->rsrc = h->lrp = h->lci = h->ltrp = h->ltci = nullptr;
    h->lval = h->ltval = nullptr;
    h->lx = h->lb = nullptr;
    h->st_tu = h->st_uniq = nullptr;
    h->st_l16 = nullptr;
    h->st_rt = h->st_ntiles = h->st_maxu = 0;
    h->st_nlist = 0;
    h->lay_ready = false;
}

int build_solver_layout(stp_csr* h, hipStream_t st) {
    if (h->lay_ready || h->gn[0] <= 0) return STP_OK;
    const int rows = h->rows, cols = h->cols;
    const int64_t nnz = h->nnz;
    const int64_t nx = h->gn[0], ny = h->gn[1], nz = h->gn[2];
    uint64_t *k0 = nullptr, *k1 = nullptr;
    int *i0 = nullptr, *i1 = nullptr, *len = nullptr;
    const int64_t nk = rows > cols ? rows : cols;
    hipError_t e = hipSuccess;
    do {
        if ((e = hipMalloc(&h->cperm, sizeof(int) * (cols ? cols : 1))) != hipSuccess) break;
        if ((e = hipMalloc(&h->rsrc, sizeof(int) * (rows ? rows : 1))) != hipSuccess) break;
        if ((e = hipMalloc(&h->lrp, sizeof(int) * ((size_t)rows + 1))) != hipSuccess) break;
        if ((e = hipMalloc(&h->lci, sizeof(int) * (nnz ? nnz : 1))) != hipSuccess) break;
        if ((e = hipMalloc(&h->lval, sizeof(float) * (nnz ? nnz : 1))) != hipSuccess) break;
        if ((e = hipMalloc(&h->ltrp, sizeof(int) * ((size_t)cols + 1))) != hipSuccess) break;
        if ((e = hipMalloc(&h->ltci, sizeof(int) * (nnz ? nnz : 1))) != hipSuccess) break;
        if ((e = hipMalloc(&h->ltval, sizeof(float) * (nnz ? nnz : 1))) != hipSuccess) break;
        if ((e = hipMalloc(&h->lx, sizeof(double) * (cols ? cols : 1))) != hipSuccess) break;
        if ((e = hipMalloc(&h->lb, sizeof(double) * (rows ? rows : 1))) != hipSuccess) break;
        if ((e = hipMalloc(&k0, sizeof(uint64_t) * (nk ? nk : 1))) != hipSuccess) break;
        if ((e = hipMalloc(&k1, sizeof(uint64_t) * (nk ? nk : 1))) != hipSuccess) break;
        if ((e = hipMalloc(&i0, sizeof(int) * (nk ? nk : 1))) != hipSuccess) break;
        if ((e = hipMalloc(&i1, sizeof(int) * (nk ? nk : 1))) != hipSuccess) break;
        if ((e = hipMalloc(&len, sizeof(int) * ((size_t)rows + 1))) != hipSuccess) break;
        // cells: Morton order (coordinates above 21 bits are truncated to 21;
        // the stable sort keeps x-major order among equal codes)
        const int64_t nmax = nx > ny ? (nx > nz ? nx : nz) : (ny > nz ? ny : nz);
        const int bc = bits_for(nmax), nb = bc > 21 ? 21 : bc, shift = bc - nb;
        if (cols > 0) {
            hipLaunchKernelGGL(cell_code_kernel, dim3(blocks(cols)), dim3(256), 0, st, nx, ny, nz, nb, shift, k0, i0);
            if ((e = hipGetLastError()) != hipSuccess) break;
            if ((e = sort_pairs(k0, k1, i0, i1, cols, nb > 0 ? 3 * nb : 1, st)) != hipSuccess) break;
            hipLaunchKernelGGL(rank_kernel, dim3(blocks(cols)), dim3(256), 0, st, cols, i1, h->cperm);
            if ((e = hipGetLastError()) != hipSuccess) break;
        }
        // rows: 6-D Morton code of their end cells
        if (rows > 0) {
            hipLaunchKernelGGL(row_key_kernel, dim3(blocks(rows)), dim3(256), 0, st, rows, h->rp, h->ci, nx, ny, nz,
                               k0, i0);
            if ((e = hipGetLastError()) != hipSuccess) break;
            if ((e = sort_pairs(k0, k1, i0, h->rsrc, rows, 60, st)) != hipSuccess) break;
        }
        hipLaunchKernelGGL(row_len_kernel, dim3(blocks((int64_t)rows + 1)), dim3(256), 0, st, rows, h->rsrc, h->rp,
                           len);
        if ((e = hipGetLastError()) != hipSuccess) break;
        if ((e = exclusive_scan_i32(len, h->lrp, rows + 1, st)) != hipSuccess) break;
        if (rows > 0) {
            hipLaunchKernelGGL(fill_kernel, dim3(blocks((int64_t)rows * 64)), dim3(256), 0, st, rows, h->rsrc, h->rp,
                               h->ci, h->val, h->cperm, h->lrp, h->lci, h->lval);
            if ((e = hipGetLastError()) != hipSuccess) break;
        }
        if ((e = csr_transpose(rows, cols, nnz, h->lrp, h->lci, h->lval, h->ltrp, h->ltci, h->ltval, st)) !=
            hipSuccess)
            break;
        h->lW = spmv_lanes_layout(rows ? (double)nnz / rows : 0.0, 0);
        h->lWT = spmv_lanes_layout(cols ? (double)nnz / cols : 0.0, 1);
        if ((e = build_staged_transpose(h, st)) != hipSuccess) break;
        e = hipStreamSynchronize(st);
    } while (0);
    (void)hipFree(k0); (void)hipFree(k1); (void)hipFree(i0); (void)hipFree(i1); (void)hipFree(len);
    if (e != hipSuccess) {
        free_solver_layout(h);
        return hip_fail(e, "solver layout");
    }
    h->lay_ready = true;
    return STP_OK;
}

template <typename T>
__global__ void permute_kernel(int n, const T* __restrict__ src, const int* __restrict__ idx, T* __restrict__ dst,
                               int scatter) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (scatter)
        dst[idx[i]] = src[i];
    else
        dst[i] = src[idx[i]];
}

template <typename T>
hipError_t permute(int n, const T* src, const int* idx, T* dst, bool scatter, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(permute_kernel<T>, dim3(blocks(n)), dim3(256), 0, st, n, src, idx, dst, scatter ? 1 : 0);
    return hipGetLastError();
}
template hipError_t permute<float>(int, const float*, const int*, float*, bool, hipStream_t);
template hipError_t permute<double>(int, const double*, const int*, double*, bool, hipStream_t);

}  // namespace stp

extern "C" {

int stp_csr_set_grid(stp_csr* h, int64_t nx, int64_t ny, int64_t nz) {
    if (!h) return stp::fail(STP_ERR_INVALID, "stp_csr_set_grid: null handle");
    if (nx == 0 && ny == 0 && nz == 0) {
        stp::free_solver_layout(h);
        h->gn[0] = h->gn[1] = h->gn[2] = 0;
        return STP_OK;
    }
    if (nx < 1 || ny < 1 || nz < 1 || nx * ny > ((int64_t)1 << 40) / nz || nx * ny * nz != h->cols)
        return stp::fail(STP_ERR_INVALID, "stp_csr_set_grid: nx * ny * nz must equal the column count");
    if (h->gn[0] != nx || h->gn[1] != ny || h->gn[2] != nz) stp::free_solver_layout(h);
    h->gn[0] = nx; h->gn[1] = ny; h->gn[2] = nz;
    return STP_OK;
}

int stp_csr_prepare_layout(stp_csr* h, void* stream) {
    if (!h) return stp::fail(STP_ERR_INVALID, "stp_csr_prepare_layout: null handle");
    return stp::build_solver_layout(h, stp::S(stream));
}

int stp_csr_staging(const stp_csr* h, int32_t* tile_rows, int64_t* list_entries) {
    if (!h || !tile_rows || !list_entries) return stp::fail(STP_ERR_INVALID, "stp_csr_staging: null argument");
    *tile_rows = h->lay_ready ? h->st_rt : -1;
    *list_entries = h->lay_ready ? h->st_nlist : 0;
    return STP_OK;
}

}  // extern "C"
