// CSR matrix (drop-in for the reference's src/cpp/sparse/csr_matrix.h:11-71).
// The canonical FP64 / int64 arrays live on the host, as the reference's
// accessors return them; the first SpMV or solve builds a device copy
// (FP32 values, int32 indices, and an explicit transpose) through the C ABI,
// and any change to the host arrays drops it.
#pragma once

#include <cstdint>
#include <vector>

#include "aligned_alloc.h"

namespace seismic {

struct SparseEntry {          // csr_matrix.h:11-15
    int64_t row;
    int64_t col;
    double value;
};

class CSRMatrix {
public:
    CSRMatrix();
    CSRMatrix(int64_t rows, int64_t cols, int64_t nnz);
    CSRMatrix(CSRMatrix&& o) noexcept;
    CSRMatrix& operator=(CSRMatrix&& o) noexcept;
    CSRMatrix(const CSRMatrix&) = delete;
    CSRMatrix& operator=(const CSRMatrix&) = delete;
    ~CSRMatrix();
    CSRMatrix clone() const;

    // entries with row or column out of range are dropped; rows are ordered
    // by column; duplicates are summed in input order (canonical CSR: the
    // reference's pad_to_alignment pass is not applied)
    void build_from_triplets(int64_t rows, int64_t cols, const std::vector<SparseEntry>& entries);
    // the reference's 2nd-order leapfrog step as a matrix on the interior grid
    // (csr_matrix.cpp:74-116)
    void build_fd_stencil_7point(int64_t nx, int64_t ny, int64_t nz, const double* velocity, double dx, double dy,
                                 double dz, double dt);
    // set the arrays directly (row_ptr: rows + 1 entries)
    void assign(int64_t rows, int64_t cols, AlignedVector<int64_t>&& row_ptr, AlignedVector<int64_t>&& col_ind,
                AlignedVector<double>&& values);

    // y = A x / A^T x on the GPU: FP64 x and accumulation, FP32 matrix values
    AlignedVector<double> spmv(const AlignedVector<double>& x) const;
    AlignedVector<double> spmv_transpose(const AlignedVector<double>& x) const;

    int64_t rows() const { return rows_; }
    int64_t cols() const { return cols_; }
    int64_t nnz() const { return nnz_; }
    const AlignedVector<int64_t>& row_ptr() const { return row_ptr_; }
    const AlignedVector<int64_t>& col_ind() const { return col_ind_; }
    const AlignedVector<double>& values() const { return values_; }

    double frobenius_norm() const;
    void scale(double alpha);
    // adds alpha to the first stored entry with col == row of each row
    void add_diagonal(double alpha);

    // CPU cache tuning in the reference (csr_matrix.cpp:195-221); kept for
    // the API, they only record the blocking
    struct RowBlock {
        int64_t row_start;
        int64_t row_end;
        int64_t nnz_count;
        int64_t pad_to_cache_line;
    };
    void optimize_cache_blocking(int64_t block_size = CACHE_LINE_SIZE / sizeof(double));
    const std::vector<RowBlock>& cache_blocks() const { return cache_blocks_; }
    void prefetch_row(int64_t) const {}

    // device copy of this matrix (an stp_csr*, built on first use)
    void* device_handle() const;

    // New: the columns are the cells of an nx x ny x nz grid (x-major, z
    // fastest); SparseOps::cgls then solves in the Morton solver layout
    // (stp_csr_set_grid).  0, 0, 0 clears it.
    void set_cell_grid(int64_t nx, int64_t ny, int64_t nz);

private:
    int64_t rows_ = 0;
    int64_t cols_ = 0;
    int64_t nnz_ = 0;
    AlignedVector<int64_t> row_ptr_;
    AlignedVector<int64_t> col_ind_;
    AlignedVector<double> values_;
    std::vector<RowBlock> cache_blocks_;
    int64_t block_size_ = 8;
    mutable void* dev_ = nullptr;
    int64_t grid_[3] = {0, 0, 0};
    void drop_device() const;
};

}  // namespace seismic
