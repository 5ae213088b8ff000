// common.hpp -- error handling and device buffers for the blqh library.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/blqh.h"

namespace blqh {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define BLQH_HIP(call)                                                                      \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            throw ::blqh::Error(BLQH_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_) + \
                                                  " (" + __FILE__ + ":" + std::to_string(__LINE__) + ")"); \
    } while (0)

#define BLQH_REQUIRE(cond, code, msg)                    \
    do {                                                 \
        if (!(cond)) throw ::blqh::Error((code), (msg)); \
    } while (0)

// Device memory of DevBuf.  Released blocks are kept and handed out again
// for the same size: sweeps free and re-allocate the same matrix sizes at
// every parameter point, hipMalloc of large blocks costs more than the kernels
// that fill them, and hipFree waits for the whole device (a solve running on
// another host thread's stream included).  A block is handed out again once
// the work queued on the releasing thread's stream before its release is
// done.  The cache holds at most a quarter of device memory and 4096 blocks
// (oldest go first) and is emptied when an allocation fails.
void* dev_alloc(size_t bytes);                 // throws Error(BLQH_ENOMEM)
void dev_free(void* p, size_t bytes) noexcept;
size_t dev_cached_bytes();                     // free for the library: counted as free memory
size_t dev_release_cached();                   // hipFree every cached block; returns the bytes released

// Owning device allocation.
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t count) {
        release();
        if (count == 0) return;
        p = static_cast<T*>(dev_alloc(count * sizeof(T)));
        n = count;
    }
    void release() {
        if (p) dev_free(p, n * sizeof(T));
        p = nullptr;
        n = 0;
    }
    size_t bytes() const { return n * sizeof(T); }
    T* get() const { return p; }
};

// The library's stream and device (one GPU per process).
struct Runtime {
    int device = -1;
    int num_cus = 0;
};
Runtime& runtime();   // initialises on first use
// The calling host thread's library stream (each thread has its own, so
// builds on one thread can overlap solves on another).  A call returns after
// its work is done, or leaves it queued on this stream only.
hipStream_t stream();

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Blocking copy ordered on the library stream: queued behind every kernel
// already on it, returns when the copy is done.  (A plain hipMemcpy runs on
// the null stream, which the non-blocking library stream does not wait for.)
inline void copy_sync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    BLQH_HIP(hipMemcpyAsync(dst, src, bytes, kind, stream()));
    BLQH_HIP(hipStreamSynchronize(stream()));
}

}  // namespace blqh
