// 64-byte aligned, move-only host vector (drop-in for the reference's
// AlignedVector, src/cpp/sparse/aligned_alloc.h:25-91).  Unlike the
// reference's, (n, value) and resize(n, value) always initialise: the
// reference skips the fill when value == T(), which leaves CG's x and
// spmv_transpose's accumulator uninitialised.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <new>
#include <utility>

namespace seismic {

constexpr size_t CACHE_LINE_SIZE = 64;

template <typename T>
class AlignedVector {
public:
    AlignedVector() = default;
    explicit AlignedVector(size_t n, T value = T()) { resize(n, value); }
    AlignedVector(const AlignedVector&) = delete;
    AlignedVector& operator=(const AlignedVector&) = delete;
    AlignedVector(AlignedVector&& o) noexcept { swap(o); }
    AlignedVector& operator=(AlignedVector&& o) noexcept {
        if (this != &o) {
            release();
            swap(o);
        }
        return *this;
    }
    ~AlignedVector() { release(); }

    void resize(size_t n, T value = T()) {
        if (n > cap_) {
            const size_t bytes = ((n * sizeof(T) + CACHE_LINE_SIZE - 1) / CACHE_LINE_SIZE) * CACHE_LINE_SIZE;
            T* p = static_cast<T*>(std::aligned_alloc(CACHE_LINE_SIZE, bytes));
            if (!p) throw std::bad_alloc();
            std::copy(data_, data_ + size_, p);
            std::free(data_);
            data_ = p;
            cap_ = n;
        }
        if (n > size_) std::fill(data_ + size_, data_ + n, value);
        size_ = n;
    }
    void assign(size_t n, T value) {
        resize(n, value);
        std::fill(data_, data_ + n, value);
    }
    AlignedVector clone() const {
        AlignedVector c(size_);
        std::copy(data_, data_ + size_, c.data_);
        return c;
    }

    T* data() { return data_; }
    const T* data() const { return data_; }
    size_t size() const { return size_; }
    bool empty() const { return size_ == 0; }
    T& operator[](size_t i) { return data_[i]; }
    const T& operator[](size_t i) const { return data_[i]; }
    T* begin() { return data_; }
    T* end() { return data_ + size_; }
    const T* begin() const { return data_; }
    const T* end() const { return data_ + size_; }

private:
    T* data_ = nullptr;
    size_t size_ = 0;
    size_t cap_ = 0;
    void release() {
        std::free(data_);
        data_ = nullptr;
        size_ = cap_ = 0;
    }
    void swap(AlignedVector& o) noexcept {
        std::swap(data_, o.data_);
        std::swap(size_, o.size_);
        std::swap(cap_, o.cap_);
    }
};

}  // namespace seismic
