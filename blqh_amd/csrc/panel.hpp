// panel.hpp -- panel-major Hamiltonian layout (see panel.hip) and the
// per-row segment product used by the panel SpMV kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "common.hpp"

namespace blqh {

constexpr int kPanelWidth = 1 << 16;

struct Panels {
    bool built = false;
    bool real = false;
    bool jagged = false;     // entry order inside each (64-row slice, panel) run: row-major or step-major
    int npan = 0;
    DevBuf<int64_t> pbase;   // [npan + 1]
    DevBuf<uint32_t> prow;   // [npan][n + 1]
    DevBuf<uint16_t> pcol;   // [nnz]
    DevBuf<double> pvr;      // [nnz] when real
    DevBuf<double2> pvc;     // [nnz] when complex
};

struct PanelView {
    int64_t n;
    int npan;
    const int64_t* pbase;
    const uint32_t* prow;
    const uint16_t* pcol;
    const double* pvr;
    const double2* pvc;
};

// number of set bits of m that belong to lanes below the calling lane
__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

inline PanelView view(const Panels& P, int64_t n) {
    return PanelView{n, P.npan, P.pbase.get(), P.prow.get(), P.pcol.get(), P.pvr.get(), P.pvc.get()};
}

}  // namespace blqh
