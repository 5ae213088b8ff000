// mtstart.hpp -- the reference's seeded start vector drawn on the device.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace blqh {

// The std::mt19937(seed) start vector on the device, bit for bit the entries
// and running norm of the host's start_entries: the calling workgroup (at
// least 312 threads) twists the 624-word state in LDS (the standard's
// recurrence in its four dependency phases: words 0..226 from the old state,
// 227..453 and 454..622 from words the previous phase wrote, 623 from words 0
// and 396), tempers, and turns draws 2i / 2i+1 into the imaginary / real part
// of entry i (u * 1.0 / (2^32 - 1) - 0.5); one lane adds |z_i|^2 in entry
// order.  *nrm2 = the sum (entries left unnormalised).
__device__ inline void mt_start_block(int64_t n, uint32_t seed, double2* __restrict__ v,
                                      double* __restrict__ nrm2) {
#pragma clang fp contract(off)
    constexpr uint32_t kUp = 0x80000000u, kLo = 0x7fffffffu, kA = 0x9908b0dfu;
    __shared__ uint32_t mt[624];
    __shared__ double2 ent[312];
    const int t = threadIdx.x;
    if (t == 0) {
        uint32_t x = seed;
        mt[0] = x;
        for (uint32_t i = 1; i < 624; ++i) {
            x = 1812433253u * (x ^ (x >> 30)) + i;
            mt[i] = x;
        }
    }
    __syncthreads();
    double s = 0.0;
    for (int64_t e0 = 0; e0 < n; e0 += 312) {
        const int lo[4] = {0, 227, 454, 623}, hi[4] = {227, 454, 623, 624};
        for (int ph = 0; ph < 4; ++ph) {
            const int i = lo[ph] + t;
            uint32_t nv = 0;
            if (i < hi[ph]) {
                const uint32_t y = (mt[i] & kUp) | (mt[i + 1 < 624 ? i + 1 : 0] & kLo);
                nv = mt[i + 397 < 624 ? i + 397 : i - 227] ^ (y >> 1) ^ ((y & 1u) ? kA : 0u);
            }
            __syncthreads();
            if (i < hi[ph]) mt[i] = nv;
            __syncthreads();
        }
        if (t < 312) {
            uint32_t z[2];
            for (int k = 0; k < 2; ++k) {
                uint32_t w = mt[2 * t + k];
                w ^= w >> 11;
                w ^= (w << 7) & 0x9d2c5680u;
                w ^= (w << 15) & 0xefc60000u;
                w ^= w >> 18;
                z[k] = w;
            }
            const double im = z[0] * 1.0 / 4294967295u - 0.5;
            const double re = z[1] * 1.0 / 4294967295u - 0.5;
            ent[t] = make_double2(re, im);
            if (e0 + t < n) v[e0 + t] = make_double2(re, im);
        }
        __syncthreads();
        if (t == 0) {
            const int cnt = (int)(n - e0 < 312 ? n - e0 : 312);
            for (int k = 0; k < cnt; ++k) {
                const double re = ent[k].x, im = ent[k].y;
                s += re * re + im * im;
            }
        }
        // ent is rewritten only after the next block's twist barriers
    }
    if (t == 0) *nrm2 = s;
}

}  // namespace blqh
