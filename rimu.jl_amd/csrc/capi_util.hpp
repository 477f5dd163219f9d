// capi_util.hpp -- host helpers shared by the C-ABI translation units: the per-thread
// last-error string, HIP error propagation and a growable device buffer.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "../../include/rimu_hip.h"

namespace rimu {

extern thread_local std::string g_last_error;

inline int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

#define HIP_TRY(expr)                                                                     \
    do {                                                                                  \
        hipError_t _e = (expr);                                                           \
        if (_e != hipSuccess)                                                             \
            return ::rimu::fail(RIMU_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(_e)); \
    } while (0)

// A device buffer that owns its memory, or (after borrow) views `count` elements of memory
// owned elsewhere: a borrowed buffer never allocates or frees, and asking it to grow past
// its size fails with hipErrorOutOfMemory.
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    bool borrowed = false;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n), borrowed(o.borrowed) { o.p = nullptr; o.n = 0; o.borrowed = false; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            n = o.n;
            borrowed = o.borrowed;
            o.p = nullptr;
            o.n = 0;
            o.borrowed = false;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    // Grows to at least `count`; a buffer that already exists grows by a quarter more than
    // asked, so sizes that creep upward from step to step do not reallocate every step.
    hipError_t alloc(size_t count) {
        if (count <= n) return hipSuccess;
        if (borrowed) return hipErrorOutOfMemory;
        const size_t want = n == 0 ? count : std::max(count, count + count / 4);
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipMalloc(&p, std::max<size_t>(want, 1) * sizeof(T));
        if (e == hipSuccess) n = want;
        return e;
    }
    void borrow(T* q, size_t count) {
        release();
        p = q;
        n = count;
        borrowed = true;
    }
    void release() {
        if (p && !borrowed) (void)hipFree(p);
        p = nullptr;
        n = 0;
        borrowed = false;
    }
};

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace rimu
