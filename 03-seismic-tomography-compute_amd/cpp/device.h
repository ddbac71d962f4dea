// Host-side helpers over the C ABI (internal to the C++ drop-in layer).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>

#include "seismic_hip.h"

namespace seismic {
namespace dev {

// throws std::runtime_error with the library's message on a failed status
inline void check(int status, const char* what) {
    if (status != STP_OK) throw std::runtime_error(std::string(what) + ": " + stp_last_error());
}

template <typename T>
T* alloc(int64_t n) {
    void* p = nullptr;
    check(stp_device_alloc(&p, static_cast<int64_t>(sizeof(T)) * (n > 0 ? n : 1)), "device allocation");
    return static_cast<T*>(p);
}

template <typename T>
void release(T*& p) {
    if (p) stp_device_free(p);
    p = nullptr;
}

template <typename T>
void upload(T* dst, const T* src, int64_t n) {
    if (n > 0) check(stp_memcpy_h2d(dst, src, static_cast<int64_t>(sizeof(T)) * n, nullptr), "host to device copy");
}

template <typename T>
void download(T* dst, const T* src, int64_t n) {
    if (n > 0) {
        check(stp_memcpy_d2h(dst, src, static_cast<int64_t>(sizeof(T)) * n, nullptr), "device to host copy");
        check(stp_stream_sync(nullptr), "stream synchronise");
    }
}

// RAII device buffer
template <typename T>
class Buffer {
public:
    explicit Buffer(int64_t n = 0) : n_(n) {
        if (n > 0) p_ = alloc<T>(n);
    }
    Buffer(const Buffer&) = delete;
    Buffer& operator=(const Buffer&) = delete;
    ~Buffer() { release(p_); }
    T* get() const { return p_; }
    int64_t size() const { return n_; }

private:
    T* p_ = nullptr;
    int64_t n_ = 0;
};

}  // namespace dev
}  // namespace seismic
