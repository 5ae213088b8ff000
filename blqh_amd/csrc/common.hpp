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
        hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            throw Error(BLQH_ENOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed: " +
                                         hipGetErrorString(e));
        }
        n = count;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    size_t bytes() const { return n * sizeof(T); }
    T* get() const { return p; }
};

// The library's stream and device (one GPU per process).
struct Runtime {
    int device = -1;
    hipStream_t stream = nullptr;
    int num_cus = 0;
};
Runtime& runtime();   // initialises on first use
hipStream_t stream();

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace blqh
