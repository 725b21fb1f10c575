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

// Nodal-block SpMV of a device CSR (csr.hip): y = mask .* (A x) (mode 0), with |A| |x| (mode 1),
// or mode 2: also partial[b] = grid block b's share of x . y, *nparts blocks (<= 4096).
// With pred non-null the launch does nothing while *pred == 0 (device-side solver control).
int csr_block_spmv(const yeti_csr *c, const double *vals, const double *x, const uint8_t *mask, double *y,
                   double *partial, int mode, int *nparts, hipStream_t st, const int32_t *pred = nullptr);
int64_t csr_nrows(const yeti_csr *c);
const int64_t *csr_rowptr_dev(const yeti_csr *c);
const int32_t *csr_colind_dev(const yeti_csr *c);

// Host-side range checks of a model (ids, spans, array lengths, load and condition targets);
// 0 or YETI_ERR_INVALID_ARGUMENT / YETI_ERR_UNSUPPORTED_ELEMENT with the message set. With
// require_solid, every patch must be a solid (U1 / U0) patch.
int validate_system_inputs(const yeti_system_inputs *in, bool require_solid);

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

// Device memory of a set-up can come from a caller-owned arena instead of hipMalloc: while an
// AllocScope is alive on a thread, DevBuf::alloc bump-allocates from its arena (256-byte
// aligned) and uploads are hipMemcpyAsync on the scope's stream. An arena with a null base only
// counts the bytes a set-up would take (no device memory is touched; set-up code skips its
// kernel launches and copies while counting()).
struct Arena {
    char *base = nullptr;   // null: count only
    size_t size = 0;
    size_t used = 0;
};

struct AllocCtx {
    Arena *arena = nullptr;
    hipStream_t stream = nullptr;
};

AllocCtx &alloc_ctx();   // per thread

struct AllocScope {
    AllocCtx saved;
    AllocScope(Arena *a, hipStream_t st) : saved(alloc_ctx()) {
        alloc_ctx().arena = a;
        alloc_ctx().stream = st;
    }
    ~AllocScope() { alloc_ctx() = saved; }
    AllocScope(const AllocScope &) = delete;
    AllocScope &operator=(const AllocScope &) = delete;
};

inline bool counting() { return alloc_ctx().arena && !alloc_ctx().arena->base; }

// Device buffer used by host-side set-up code: owns hipMalloc memory, or borrows arena memory.
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    bool owned = false;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p && owned) (void)hipFree(p);
        p = nullptr;
        n = 0;
        owned = false;
    }
    hipError_t alloc(size_t count) {
        release();
        n = count;
        if (count == 0) return hipSuccess;
        const size_t bytes = count * sizeof(T);
        if (Arena *a = alloc_ctx().arena) {
            const size_t off = (a->used + 255) & ~size_t(255);
            if (a->base && (off > a->size || bytes > a->size - off)) return hipErrorOutOfMemory;
            a->used = off + bytes;
            // counting: a distinct non-null placeholder that is never dereferenced
            p = a->base ? reinterpret_cast<T *>(a->base + off) : reinterpret_cast<T *>(uintptr_t(off + 256));
            return hipSuccess;
        }
        owned = true;
        return hipMalloc((void **)&p, bytes);
    }
    hipError_t upload(const T *h, size_t count) {
        hipError_t e = alloc(count);
        if (e != hipSuccess || count == 0 || counting()) return e;
        if (alloc_ctx().arena) {
            // the source may be a temporary of the set-up: wait for this stream only
            e = hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, alloc_ctx().stream);
            return e == hipSuccess ? hipStreamSynchronize(alloc_ctx().stream) : e;
        }
        return hipMemcpy(p, h, count * sizeof(T), hipMemcpyHostToDevice);
    }
};

}  // namespace yeti
