// CSRMatrix (reference interface: src/cpp/sparse/csr_matrix.h:17-71).  Assembly
// and SpMV run on the GPU through the C ABI; the host keeps the canonical
// FP64 / int64 arrays the accessors return.
#include "sparse/csr_matrix.h"

#include <cmath>
#include <stdexcept>
#include <utility>

#include "device.h"

namespace seismic {

CSRMatrix::CSRMatrix() : row_ptr_(1, 0) {}

CSRMatrix::CSRMatrix(int64_t rows, int64_t cols, int64_t nnz)
    : rows_(rows), cols_(cols), nnz_(nnz), row_ptr_(static_cast<size_t>(rows + 1), 0),
      col_ind_(static_cast<size_t>(nnz), 0), values_(static_cast<size_t>(nnz), 0.0) {}

CSRMatrix::CSRMatrix(CSRMatrix&& o) noexcept { *this = std::move(o); }

CSRMatrix& CSRMatrix::operator=(CSRMatrix&& o) noexcept {
    if (this != &o) {
        drop_device();
        rows_ = o.rows_;
        cols_ = o.cols_;
        nnz_ = o.nnz_;
        row_ptr_ = std::move(o.row_ptr_);
        col_ind_ = std::move(o.col_ind_);
        values_ = std::move(o.values_);
        cache_blocks_ = std::move(o.cache_blocks_);
        block_size_ = o.block_size_;
        for (int k = 0; k < 3; ++k) grid_[k] = o.grid_[k];
        dev_ = o.dev_;
        o.dev_ = nullptr;
        o.rows_ = o.cols_ = o.nnz_ = 0;
    }
    return *this;
}

CSRMatrix::~CSRMatrix() { drop_device(); }

CSRMatrix CSRMatrix::clone() const {
    CSRMatrix c;
    c.rows_ = rows_;
    c.cols_ = cols_;
    c.nnz_ = nnz_;
    c.row_ptr_ = row_ptr_.clone();
    c.col_ind_ = col_ind_.clone();
    c.values_ = values_.clone();
    c.cache_blocks_ = cache_blocks_;
    c.block_size_ = block_size_;
    for (int k = 0; k < 3; ++k) c.grid_[k] = grid_[k];
    return c;
}

void CSRMatrix::drop_device() const {
    if (dev_) stp_csr_destroy(static_cast<stp_csr*>(dev_));
    dev_ = nullptr;
}

void CSRMatrix::assign(int64_t rows, int64_t cols, AlignedVector<int64_t>&& rp, AlignedVector<int64_t>&& ci,
                       AlignedVector<double>&& v) {
    if (rows < 0 || cols < 0 || rp.size() != static_cast<size_t>(rows + 1) || ci.size() != v.size())
        throw std::invalid_argument("CSRMatrix::assign: inconsistent arrays");
    drop_device();
    if (cols != cols_) grid_[0] = grid_[1] = grid_[2] = 0;   // a grid of the old column count
    rows_ = rows;
    cols_ = cols;
    nnz_ = static_cast<int64_t>(ci.size());
    row_ptr_ = std::move(rp);
    col_ind_ = std::move(ci);
    values_ = std::move(v);
}

// Device assembly (stp_csr_assemble) of triplets already on the device.
static void assemble_device(int64_t rows, int64_t cols, int64_t n, const int64_t* r, const int64_t* c,
                            const double* v, AlignedVector<int64_t>& rp_out, AlignedVector<int64_t>& ci_out,
                            AlignedVector<double>& val_out) {
    dev::Buffer<int64_t> rp(rows + 1), ci(n);
    dev::Buffer<double> val(n);
    int64_t nnz = 0;
    dev::check(stp_csr_assemble(rows, cols, n, r, c, v, rp.get(), ci.get(), val.get(), &nnz, nullptr),
               "CSRMatrix assembly");
    rp_out = AlignedVector<int64_t>(static_cast<size_t>(rows + 1));
    ci_out = AlignedVector<int64_t>(static_cast<size_t>(nnz));
    val_out = AlignedVector<double>(static_cast<size_t>(nnz));
    dev::download(rp_out.data(), rp.get(), rows + 1);
    dev::download(ci_out.data(), ci.get(), nnz);
    dev::download(val_out.data(), val.get(), nnz);
}

void CSRMatrix::build_from_triplets(int64_t rows, int64_t cols, const std::vector<SparseEntry>& entries) {
    if (rows < 0 || cols < 0) throw std::invalid_argument("CSRMatrix::build_from_triplets: negative shape");
    const int64_t n = static_cast<int64_t>(entries.size());
    AlignedVector<int64_t> rp, ci;
    AlignedVector<double> val;
    if (n == 0) {
        rp = AlignedVector<int64_t>(static_cast<size_t>(rows + 1), 0);
    } else {
        std::vector<int64_t> r(entries.size()), c(entries.size());
        std::vector<double> v(entries.size());
        for (size_t k = 0; k < entries.size(); ++k) {
            r[k] = entries[k].row;
            c[k] = entries[k].col;
            v[k] = entries[k].value;
        }
        dev::Buffer<int64_t> dr(n), dc(n);
        dev::Buffer<double> dv(n);
        dev::upload(dr.get(), r.data(), n);
        dev::upload(dc.get(), c.data(), n);
        dev::upload(dv.get(), v.data(), n);
        assemble_device(rows, cols, n, dr.get(), dc.get(), dv.get(), rp, ci, val);
    }
    assign(rows, cols, std::move(rp), std::move(ci), std::move(val));
}

void CSRMatrix::build_fd_stencil_7point(int64_t nx, int64_t ny, int64_t nz, const double* velocity, double dx,
                                        double dy, double dz, double dt) {
    const int64_t V = nx * ny * nz;
    if (V <= 0) throw std::invalid_argument("CSRMatrix::build_fd_stencil_7point: empty grid");
    dev::Buffer<double> vel(V), dv(7 * V);
    dev::Buffer<int64_t> dr(7 * V), dc(7 * V);
    dev::upload(vel.get(), velocity, V);
    dev::check(stp_csr64_fd7_triplets(nx, ny, nz, vel.get(), dx, dy, dz, dt, dr.get(), dc.get(), dv.get(), nullptr),
               "build_fd_stencil_7point");
    AlignedVector<int64_t> rp, ci;
    AlignedVector<double> val;
    assemble_device(V, V, 7 * V, dr.get(), dc.get(), dv.get(), rp, ci, val);
    assign(V, V, std::move(rp), std::move(ci), std::move(val));
}

void* CSRMatrix::device_handle() const {
    if (dev_) return dev_;
    if (rows_ > 0x7ffffffe || cols_ > 0x7ffffffe || nnz_ > 0x7fffffff)
        throw std::length_error("CSRMatrix: too large for the device format (int32 indices)");
    dev::Buffer<int64_t> rp(rows_ + 1), ci(nnz_);
    dev::Buffer<double> v(nnz_);
    dev::Buffer<int32_t> rp32(rows_ + 1), ci32(nnz_);
    dev::Buffer<float> v32(nnz_);
    dev::upload(rp.get(), row_ptr_.data(), rows_ + 1);
    dev::upload(ci.get(), col_ind_.data(), nnz_);
    dev::upload(v.get(), values_.data(), nnz_);
    dev::check(stp_csr64_to_f32(rows_, nnz_, rp.get(), ci.get(), v.get(), rp32.get(), ci32.get(), v32.get(), nullptr),
               "CSRMatrix device copy");
    stp_csr* h = nullptr;
    dev::check(stp_csr_create(rows_, cols_, nnz_, rp32.get(), ci32.get(), v32.get(), &h, nullptr),
               "CSRMatrix device copy");
    // the FP64 values serve SparseOps' CG, which the reference runs in FP64
    const int rc = stp_csr_set_values_f64(h, v.get(), nullptr);
    if (rc != STP_OK) {
        stp_csr_destroy(h);
        dev::check(rc, "CSRMatrix device copy");
    }
    dev::check(stp_stream_sync(nullptr), "CSRMatrix device copy");
    dev_ = h;
    if (grid_[0] > 0) dev::check(stp_csr_set_grid(h, grid_[0], grid_[1], grid_[2]), "CSRMatrix::set_cell_grid");
    return dev_;
}

void CSRMatrix::set_cell_grid(int64_t nx, int64_t ny, int64_t nz) {
    const bool clear = nx == 0 && ny == 0 && nz == 0;
    if (!clear && (nx < 1 || ny < 1 || nz < 1 || nx * ny * nz != cols_))
        throw std::invalid_argument("CSRMatrix::set_cell_grid: nx * ny * nz must equal the column count");
    grid_[0] = nx;
    grid_[1] = ny;
    grid_[2] = nz;
    if (dev_) dev::check(stp_csr_set_grid(static_cast<stp_csr*>(dev_), nx, ny, nz), "CSRMatrix::set_cell_grid");
}

static AlignedVector<double> spmv_dev(const CSRMatrix& A, const AlignedVector<double>& x, int transpose) {
    const int64_t nin = transpose ? A.rows() : A.cols();
    const int64_t nout = transpose ? A.cols() : A.rows();
    if (static_cast<int64_t>(x.size()) < nin) throw std::invalid_argument("CSRMatrix::spmv: x is too short");
    AlignedVector<double> y(static_cast<size_t>(nout), 0.0);
    if (nout == 0) return y;
    dev::Buffer<double> dx(nin), dy(nout);
    dev::upload(dx.get(), x.data(), nin);
    dev::check(stp_csr_spmv_f64(static_cast<const stp_csr*>(A.device_handle()), transpose, dx.get(), dy.get(),
                                nullptr),
               "CSRMatrix::spmv");
    dev::download(y.data(), dy.get(), nout);
    return y;
}

AlignedVector<double> CSRMatrix::spmv(const AlignedVector<double>& x) const { return spmv_dev(*this, x, 0); }
AlignedVector<double> CSRMatrix::spmv_transpose(const AlignedVector<double>& x) const { return spmv_dev(*this, x, 1); }

double CSRMatrix::frobenius_norm() const {
    double s = 0.0;
    for (int64_t k = 0; k < nnz_; ++k) s += values_[k] * values_[k];
    return std::sqrt(s);
}

void CSRMatrix::scale(double alpha) {
    for (int64_t k = 0; k < nnz_; ++k) values_[k] *= alpha;
    drop_device();
}

void CSRMatrix::add_diagonal(double alpha) {
    for (int64_t r = 0; r < rows_; ++r)
        for (int64_t k = row_ptr_[r]; k < row_ptr_[r + 1]; ++k)
            if (col_ind_[k] == r) {
                values_[k] += alpha;
                break;
            }
    drop_device();
}

void CSRMatrix::optimize_cache_blocking(int64_t block_size) {
    block_size_ = block_size > 0 ? block_size : 8;
    cache_blocks_.clear();
    for (int64_t r0 = 0; r0 < rows_; r0 += block_size_) {
        const int64_t r1 = std::min(rows_, r0 + block_size_);
        const int64_t nz = row_ptr_[r1] - row_ptr_[r0];
        cache_blocks_.push_back(RowBlock{r0, r1, nz, 0});
    }
}

}  // namespace seismic
