// common.hpp -- shared helpers of libyeti_hip.so (error reporting, HIP checks, MFMA types).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/yeti_amd.h"

namespace yeti {

// Per-thread message of the last failing call (read through yeti_last_error()).
std::string &last_error();

int set_error(int code, const char *fmt, ...);

#define YETI_HIP(call)                                                                    \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            return ::yeti::set_error(YETI_ERR_DEVICE, "%s failed: %s (%s:%d)", #call,     \
                                     hipGetErrorString(e_), __FILE__, __LINE__);          \
    } while (0)

typedef double f64x4 __attribute__((ext_vector_type(4)));

// Gauss-Legendre rule passed to kernels by value (orders 1..9, as YETI's table).
constexpr int MAX_NG = 9;
struct GaussRule {
    double x[MAX_NG];
    double w[MAX_NG];
};

// RAII device buffer (hipMalloc / hipFree) used by host-side set-up code.
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    hipError_t alloc(size_t count) {
        release();
        n = count;
        if (count == 0) return hipSuccess;
        return hipMalloc((void **)&p, count * sizeof(T));
    }
    hipError_t upload(const T *h, size_t count) {
        hipError_t e = alloc(count);
        if (e != hipSuccess || count == 0) return e;
        return hipMemcpy(p, h, count * sizeof(T), hipMemcpyHostToDevice);
    }
};

}  // namespace yeti
