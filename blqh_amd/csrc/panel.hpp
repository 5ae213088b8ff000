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

// One matrix of a panel relayout (panel.hip): its panel arrays and a scratch
// copy of the entries.
struct RelayoutJob {
    int64_t n, nnz;
    int32_t npan, real;
    const int64_t* pbase;
    const uint32_t* prow;
    uint16_t* col;
    double* vr;
    double2* vc;
    uint16_t* tcol;
    double* tvr;
    double2* tvc;
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

// Row sums of one (64-row slice, panel) run in step-major order: step t
// holds entry t of every row with more than t entries, rows in order.  Lane l
// owns row l (`len` entries); at step t the active lanes read consecutive
// positions off_t + (active lanes below l).  cr / vr / vc point at the run's
// first entry, xp at the panel's slice of x.  Four steps per iteration, so
// four loads per stream and four x gathers are in flight per lane.
template <bool REAL>
__device__ __forceinline__ void jag_row_sum(const uint16_t* __restrict__ cr, const double* __restrict__ vr,
                                            const double2* __restrict__ vc, const double2* __restrict__ xp,
                                            uint32_t len, double& ar, double& ai) {
    typedef double d2v_t __attribute__((ext_vector_type(2)));
    uint32_t off = 0;
    for (uint32_t t = 0;; t += 4) {
        uint64_t m[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) m[u] = __ballot(t + u < len);
        if (m[0] == 0) break;
        uint32_t q[4];
        q[0] = off + lanes_below(m[0]);
        uint32_t o = off + (uint32_t)__popcll(m[0]);
#pragma unroll
        for (int u = 1; u < 4; ++u) {
            q[u] = o + lanes_below(m[u]);
            o += (uint32_t)__popcll(m[u]);
        }
        off = o;
        int c[4];
        double2 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const bool act = t + u < len;
            c[u] = act ? (int)__builtin_nontemporal_load(&cr[q[u]]) : 0;
            if (REAL) {
                v[u] = make_double2(act ? __builtin_nontemporal_load(&vr[q[u]]) : 0.0, 0.0);
            } else if (act) {
                const d2v_t w = __builtin_nontemporal_load(reinterpret_cast<const d2v_t*>(&vc[q[u]]));
                v[u] = make_double2(w.x, w.y);
            } else {
                v[u] = make_double2(0.0, 0.0);
            }
        }
        double2 xv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) xv[u] = xp[c[u]];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (REAL) {
                ar += v[u].x * xv[u].x;
                ai += v[u].x * xv[u].y;
            } else {
                ar += v[u].x * xv[u].x - v[u].y * xv[u].y;
                ai += v[u].x * xv[u].y + v[u].y * xv[u].x;
            }
        }
    }
}

inline PanelView view(const Panels& P, int64_t n) {
    return PanelView{n, P.npan, P.pbase.get(), P.prow.get(), P.pcol.get(), P.pvr.get(), P.pvc.get()};
}

}  // namespace blqh
