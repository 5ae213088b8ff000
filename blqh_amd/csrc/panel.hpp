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
    bool greedy = false;     // step-major with the greedy step order (set_panel_layout), else the plain one
    int rows = 1;            // step-major: rows per lane (1, or 2 / 4 / 8 for SpMV kind 5)
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

// Column tiles (SpMV kind 4): the columns are cut into tiles of kTileCols
// (an x tile of 128 KiB, staged in LDS), the rows into G blocks of SB 64-row
// slices, one block per workgroup of kTileWaves wavefronts; wave w of block b
// owns the block's slices w, w + 16, ... (R of them) and reads one contiguous
// stream: for every tile t, for every owned slice r, the (slice, tile) run in
// step-major order (see jag_row_sum).  len[wave][t][r][lane] holds each row's
// entry count in the run (u8; the layout is not built when a count exceeds
// 255), wstart[wave] the stream's first entry.
constexpr int kTileCols = 8192;
constexpr int kTileWaves = 16;
constexpr int kTileMaxR = 16;

struct Tiles {
    bool built = false;
    bool tried = false;   // build attempted (it may be unavailable for this matrix)
    bool real = false;
    int G = 0, SB = 0, R = 0, ntiles = 0;
    DevBuf<uint8_t> len;     // [G * kTileWaves][ntiles][R][64]
    DevBuf<int64_t> wstart;  // [G * kTileWaves]
    DevBuf<uint16_t> col;    // [nnz] column - tile base
    DevBuf<double> vr;       // [nnz] when real
    DevBuf<double2> vc;      // [nnz] when complex
};

struct TileView {
    int64_t n;
    int G, SB, R, ntiles;
    const uint8_t* len;
    const int64_t* wstart;
    const uint16_t* col;
    const double* vr;
    const double2* vc;
};

inline TileView tview(const Tiles& T, int64_t n) {
    return TileView{n, T.G, T.SB, T.R, T.ntiles, T.len.get(), T.wstart.get(), T.col.get(), T.vr.get(), T.vc.get()};
}

// Row sums of one (64-row slice, panel) run in step-major order: step t
// holds entry t of every row with more than t entries, rows in order.  Lane l
// owns row l (`len` entries); at step t the active lanes read consecutive
// positions off_t + (active lanes below l).  cr / vr / vc point at the run's
// first entry, xp at the panel's slice of x.  Four steps per iteration, so
// four loads per stream and four x gathers are in flight per lane.
template <bool REAL>
__device__ __forceinline__ uint32_t jag_row_sum(const uint16_t* __restrict__ cr, const double* __restrict__ vr,
                                                const double2* __restrict__ vc, const double2* __restrict__ xp,
                                                uint32_t len, double& ar, double& ai) {
    typedef double d2v_t __attribute__((ext_vector_type(2)));
    // Software pipeline over groups of four steps: the column and value loads
    // of group g+1 are issued before group g's gathers and products, so a
    // wavefront always has the next group's loads in flight.
    uint32_t off = 0;
    auto issue = [&](uint32_t t, int* c, double2* v, bool& any) {
        uint64_t m[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) m[u] = __ballot(t + u < len);
        any = m[0] != 0;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint32_t q = off + lanes_below(m[u]);
            off += (uint32_t)__popcll(m[u]);
            const bool act = t + u < len;
            c[u] = act ? (int)__builtin_nontemporal_load(&cr[q]) : 0;
            if (REAL) {
                v[u] = make_double2(act ? __builtin_nontemporal_load(&vr[q]) : 0.0, 0.0);
            } else if (act) {
                const d2v_t w = __builtin_nontemporal_load(reinterpret_cast<const d2v_t*>(&vc[q]));
                v[u] = make_double2(w.x, w.y);
            } else {
                v[u] = make_double2(0.0, 0.0);
            }
        }
    };
    int c0[4], c1[4];
    double2 v0[4], v1[4];
    bool any0, any1;
    issue(0, c0, v0, any0);
    for (uint32_t t = 4; any0; t += 4) {
        issue(t, c1, v1, any1);  // next group in flight
        double2 xv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) xv[u] = xp[c0[u]];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (REAL) {
                ar += v0[u].x * xv[u].x;
                ai += v0[u].x * xv[u].y;
            } else {
                ar += v0[u].x * xv[u].x - v0[u].y * xv[u].y;
                ai += v0[u].x * xv[u].y + v0[u].y * xv[u].x;
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            c0[u] = c1[u];
            v0[u] = v1[u];
        }
        any0 = any1;
    }
    return off;  // entries of the whole run (wave-uniform)
}

// jag_row_sum split over NP wavefronts: wave `par` (0 <= par < NP) sums the
// step groups g (four steps each) with g mod NP == par and only counts the
// entries of the others, so each lane's chain of dependent loads is 1/NP as
// long.  The row sum is the NP waves' partial sums added by the caller in
// wave order.
template <bool REAL, int NP>
__device__ __forceinline__ void jag_row_sum_part(const uint16_t* __restrict__ cr, const double* __restrict__ vr,
                                                 const double2* __restrict__ vc, const double2* __restrict__ xp,
                                                 uint32_t len, int par, double& ar, double& ai) {
    typedef double d2v_t __attribute__((ext_vector_type(2)));
    uint32_t off = 0;
    auto issue = [&](uint32_t t, int* c, double2* v, bool& any) {
        uint64_t m[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) m[u] = __ballot(t + u < len);
        any = m[0] != 0;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint32_t q = off + lanes_below(m[u]);
            off += (uint32_t)__popcll(m[u]);
            const bool act = t + u < len;
            c[u] = act ? (int)__builtin_nontemporal_load(&cr[q]) : 0;
            if (REAL) {
                v[u] = make_double2(act ? __builtin_nontemporal_load(&vr[q]) : 0.0, 0.0);
            } else if (act) {
                const d2v_t w = __builtin_nontemporal_load(reinterpret_cast<const d2v_t*>(&vc[q]));
                v[u] = make_double2(w.x, w.y);
            } else {
                v[u] = make_double2(0.0, 0.0);
            }
        }
    };
    auto skip = [&](uint32_t t) {  // another wave's group: advance the offset only
#pragma unroll
        for (int u = 0; u < 4; ++u) off += (uint32_t)__popcll(__ballot(t + u < len));
    };
    int c0[4], c1[4];
    double2 v0[4], v1[4];
    bool any0, any1;
    for (int g = 0; g < par; ++g) skip(4u * g);
    issue(4u * par, c0, v0, any0);
    for (uint32_t t = 4u * par + 4u * NP; any0; t += 4u * NP) {
#pragma unroll
        for (int g = NP - 1; g > 0; --g) skip(t - 4u * g);
        issue(t, c1, v1, any1);  // next own group in flight
        double2 xv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) xv[u] = xp[c0[u]];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (REAL) {
                ar += v0[u].x * xv[u].x;
                ai += v0[u].x * xv[u].y;
            } else {
                ar += v0[u].x * xv[u].x - v0[u].y * xv[u].y;
                ai += v0[u].x * xv[u].y + v0[u].y * xv[u].x;
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            c0[u] = c1[u];
            v0[u] = v1[u];
        }
        any0 = any1;
    }
}

// The same over a lane stream that covers R rows one after the other (SpMV
// kind 5, panel runs with R rows per lane): cum[q] = the lane's entries in
// rows 0..q of its stream.  The running sum is added to acc[q * 64] (this
// lane's LDS slot of row q) when the stream passes the end of row q; the row
// boundaries stay in registers (static indices, selected by compare).
template <bool REAL, int R>
__device__ __forceinline__ void jag_rows_sum(const uint16_t* __restrict__ cr, const double* __restrict__ vr,
                                             const double2* __restrict__ vc, const double2* __restrict__ xp,
                                             const uint32_t (&cum)[R], double2* acc) {
    typedef double d2v_t __attribute__((ext_vector_type(2)));
    const uint32_t len = cum[R - 1];
    int cur = 0;
    uint32_t bnd = cum[0];
    double ar = 0.0, ai = 0.0;
    uint32_t off = 0;
    auto issue = [&](uint32_t t, int* c, double2* v, bool& any) {
        uint64_t m[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) m[u] = __ballot(t + u < len);
        any = m[0] != 0;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint32_t q = off + lanes_below(m[u]);
            off += (uint32_t)__popcll(m[u]);
            const bool act = t + u < len;
            c[u] = act ? (int)__builtin_nontemporal_load(&cr[q]) : 0;
            if (REAL) {
                v[u] = make_double2(act ? __builtin_nontemporal_load(&vr[q]) : 0.0, 0.0);
            } else if (act) {
                const d2v_t w = __builtin_nontemporal_load(reinterpret_cast<const d2v_t*>(&vc[q]));
                v[u] = make_double2(w.x, w.y);
            } else {
                v[u] = make_double2(0.0, 0.0);
            }
        }
    };
    int c0[4], c1[4];
    double2 v0[4], v1[4];
    bool any0, any1;
    issue(0, c0, v0, any0);
    for (uint32_t t = 4; any0; t += 4) {
        issue(t, c1, v1, any1);  // next group in flight
        double2 xv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) xv[u] = xp[c0[u]];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint32_t tt = t - 4 + u;
            if (tt < len) {
                while (tt >= bnd && cur < R - 1) {  // past the end of row cur (rows without entries skipped)
                    double2 a = acc[cur * 64];
                    a.x += ar;
                    a.y += ai;
                    acc[cur * 64] = a;
                    ar = 0.0;
                    ai = 0.0;
                    ++cur;
#pragma unroll
                    for (int q = 0; q < R; ++q)
                        if (q == cur) bnd = cum[q];
                }
                if (REAL) {
                    ar += v0[u].x * xv[u].x;
                    ai += v0[u].x * xv[u].y;
                } else {
                    ar += v0[u].x * xv[u].x - v0[u].y * xv[u].y;
                    ai += v0[u].x * xv[u].y + v0[u].y * xv[u].x;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            c0[u] = c1[u];
            v0[u] = v1[u];
        }
        any0 = any1;
    }
    if (len) {
        double2 a = acc[cur * 64];
        a.x += ar;
        a.y += ai;
        acc[cur * 64] = a;
    }
}

inline PanelView view(const Panels& P, int64_t n) {
    return PanelView{n, P.npan, P.pbase.get(), P.prow.get(), P.pcol.get(), P.pvr.get(), P.pvc.get()};
}

}  // namespace blqh
