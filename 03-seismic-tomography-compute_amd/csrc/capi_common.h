// Shared error plumbing of the C ABI implementation files.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "seismic_hip.h"

namespace stp {
std::string& last_error();

inline int fail(int code, const std::string& msg) {
    last_error() = msg;
    return code;
}

inline int hip_fail(hipError_t e, const char* where) {
    last_error() = std::string(where) + ": " + hipGetErrorString(e);
    return STP_ERR_HIP;
}

inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }
}  // namespace stp

#define STP_HIP(call)                                           \
    do {                                                        \
        hipError_t _e = (call);                                 \
        if (_e != hipSuccess) return ::stp::hip_fail(_e, #call); \
    } while (0)
