// common.hpp -- error plumbing, device buffers and launch helpers shared by the
// EROT HIP kernels and the host drivers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace regot {

// Errors cross the C-ABI as a negative status plus a thread-local message.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

enum Status : int {
    OK = 0,
    E_INVALID = -1,   // bad argument (sizes, dtype, null pointer)
    E_HIP = -2,       // HIP runtime failure
    E_NOMEM = -3,     // device allocation failed
    E_STATE = -4,     // operation needs state that is not there (e.g. no Hessian yet)
    E_INTERNAL = -5,
    E_SOLVER = -6,    // a solver stopped with an error (e.g. a failed line search)
};

[[noreturn]] inline void fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    throw Error(code, buf);
}

#define HIP_CHECK(expr)                                                               \
    do {                                                                              \
        hipError_t _e = (expr);                                                       \
        if (_e != hipSuccess)                                                         \
            ::regot::fail(::regot::E_HIP, "%s failed: %s (%s:%d)", #expr,             \
                          hipGetErrorString(_e), __FILE__, __LINE__);                 \
    } while (0)

#define KERNEL_CHECK() HIP_CHECK(hipGetLastError())

// Owning device allocation (hipMalloc), released on destruction.
template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    explicit DBuf(size_t count) { alloc(count); }
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DBuf& operator=(DBuf&& o) noexcept
    {
        if (this != &o) {
            release();
            p = o.p; n = o.n;
            o.p = nullptr; o.n = 0;
        }
        return *this;
    }
    ~DBuf() { release(); }
    void alloc(size_t count)
    {
        release();
        if (count == 0) count = 1;
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            fail(E_NOMEM, "hipMalloc of %zu bytes failed: %s", count * sizeof(T),
                 hipGetErrorString(e));
        }
        n = count;
    }
    // Grow-only reallocation; contents are not preserved.
    void ensure(size_t count)
    {
        if (count > n || p == nullptr) alloc(count);
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    operator T*() const { return p; }
};

// Pinned host staging memory (hipHostMalloc), released on destruction.
template <typename T>
struct HBuf {
    T* p = nullptr;
    size_t n = 0;
    HBuf() = default;
    HBuf(const HBuf&) = delete;
    HBuf& operator=(const HBuf&) = delete;
    ~HBuf() { release(); }
    void ensure(size_t count)
    {
        if (count <= n && p) return;
        release();
        if (count == 0) count = 1;
        HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&p), count * sizeof(T),
                                hipHostMallocDefault));
        n = count;
    }
    void release()
    {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
    }
};

// Pinned, device-mapped host array: kernels write it, the host reads it after
// the writing work has completed (stream order + event).
template <typename T>
struct MappedRing {
    T* h = nullptr;
    T* d = nullptr;
    int cap = 0;
    MappedRing() = default;
    MappedRing(const MappedRing&) = delete;
    MappedRing& operator=(const MappedRing&) = delete;
    ~MappedRing() { release(); }
    void ensure(int c)
    {
        if (h && c <= cap) return;
        release();
        HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&h), sizeof(T) * c, hipHostMallocMapped));
        HIP_CHECK(hipHostGetDevicePointer(reinterpret_cast<void**>(&d), h, 0));
        cap = c;
    }
    void release()
    {
        if (h) (void)hipHostFree(h);
        h = nullptr;
        d = nullptr;
        cap = 0;
    }
};

inline int ceil_div(long a, long b) { return static_cast<int>((a + b - 1) / b); }

}  // namespace regot
