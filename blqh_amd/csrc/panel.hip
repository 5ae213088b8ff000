// panel.hip -- panel-major copy of the Hamiltonian for the SpMV hot loop.
//
// The CSR rows of a Hamiltonian touch ~350 columns spread over the whole
// vector, so a row-major SpMV gathers x from everywhere: at 2.3M states the
// 37 MB vector falls out of L2 and every 16-byte gather pulls a full line.
// Here the columns are cut into panels of 65536 (x slice of 1 MB for complex
// vectors, which stays resident in each XCD's 4 MB L2).  Panel p stores, for
// every row, that row's entries with p*65536 <= col < (p+1)*65536, rows in
// order, so each panel is one sequential stream:
//   pbase[p]        first entry of panel p
//   prow[p][r]      offset of row r's segment inside panel p (uint32, n+1 per panel)
//   pcol[k]         col - p*65536 (uint16)
//   pvr[k] / pvc[k] value (real double when the matrix is real, else double2)
// The SpMV kernel is persistent: workgroup b owns a contiguous row range and
// keeps those rows' partial sums in LDS while it walks the panels in order, so
// all workgroups sweep the same panel at about the same time and y is written
// once.  The matrix is stored real when every imaginary part is below 1e-14
// of the largest entry (the K = 0 sectors are exactly real; K = nphi/2 sectors
// carry sin(pi n) rounding residue of 1e-16).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "engine.hpp"
#include "panel.hpp"

namespace {

using namespace blqh;

constexpr int kPanelShift = 16;
constexpr int kBlk = 256;

// per row: entries per panel -> cnt[p * n + row]
__global__ void k_panel_count(int64_t n, int npan, const int64_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                              int32_t* __restrict__ cnt) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        int p = 0, c = 0;
        for (int64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) {
            const int q = col[k] >> kPanelShift;
            while (p < q) { cnt[(int64_t)p * n + r] = c; c = 0; ++p; }
            ++c;
        }
        while (p < npan) { cnt[(int64_t)p * n + r] = c; c = 0; ++p; }
    }
}

// off = exclusive scan of cnt over (panel, row): global destination of each segment
__global__ void k_panel_scatter(int64_t n, const int64_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                const double2* __restrict__ val, const int64_t* __restrict__ off, int real,
                                uint16_t* __restrict__ pcol, double* __restrict__ pvr, double2* __restrict__ pvc) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        int p = -1;
        int64_t dst = 0;
        for (int64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) {
            const int q = col[k] >> kPanelShift;
            if (q != p) { p = q; dst = off[(int64_t)q * n + r]; }
            pcol[dst] = (uint16_t)(col[k] & (kPanelWidth - 1));
            if (real) pvr[dst] = val[k].x; else pvc[dst] = val[k];
            ++dst;
        }
    }
}

// The row-major panel arrays again from the CSR, in place (columns ascending
// in every row, as built): run offsets from pbase / prow.
__global__ void k_panel_rescatter(int64_t n, const int64_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                  const double2* __restrict__ val, const int64_t* __restrict__ pbase,
                                  const uint32_t* __restrict__ prow, int real, uint16_t* __restrict__ pcol,
                                  double* __restrict__ pvr, double2* __restrict__ pvc) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        int p = -1;
        int64_t dst = 0;
        for (int64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) {
            const int q = col[k] >> kPanelShift;
            if (q != p) { p = q; dst = pbase[q] + prow[(int64_t)q * (n + 1) + r]; }
            pcol[dst] = (uint16_t)(col[k] & (kPanelWidth - 1));
            if (real) pvr[dst] = val[k].x; else pvc[dst] = val[k];
            ++dst;
        }
    }
}

// prow[p][r] = off[p*n + r] - pbase[p] for r <= n
__global__ void k_panel_rows(int64_t n, int npan, const int64_t* __restrict__ off, int64_t total,
                             uint32_t* __restrict__ prow, int64_t* __restrict__ pbase) {
    const int64_t m = (int64_t)npan * (n + 1);
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < m; t += (int64_t)gridDim.x * blockDim.x) {
        const int p = (int)(t / (n + 1));
        const int64_t r = t - (int64_t)p * (n + 1);
        const int64_t base = off[(int64_t)p * n];
        const int64_t g = (int64_t)p * n + r;  // r == n reads the next panel's first offset (or the total)
        const int64_t v = g < (int64_t)npan * n ? off[g] : total;
        prow[t] = (uint32_t)(v - base);
        if (r == 0) pbase[p] = base;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) pbase[npan] = total;
}

// max |re| and max |im| of the values (two doubles via atomicMax on the bit
// pattern of non-negative doubles, one atomic pair per workgroup)
__global__ void __launch_bounds__(kBlk) k_absmax(int64_t nnz, const double2* __restrict__ val,
                                                 unsigned long long* __restrict__ out) {
    __shared__ double sr[kBlk / 64], si[kBlk / 64];
    double mr = 0, mi = 0;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nnz; k += (int64_t)gridDim.x * blockDim.x) {
        mr = fmax(mr, fabs(val[k].x));
        mi = fmax(mi, fabs(val[k].y));
    }
    for (int o = 32; o > 0; o >>= 1) {
        mr = fmax(mr, __shfl_xor(mr, o, 64));
        mi = fmax(mi, __shfl_xor(mi, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        sr[threadIdx.x >> 6] = mr;
        si[threadIdx.x >> 6] = mi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBlk / 64; ++w) {
            mr = fmax(mr, sr[w]);
            mi = fmax(mi, si[w]);
        }
        atomicMax(&out[0], (unsigned long long)__double_as_longlong(mr));
        atomicMax(&out[1], (unsigned long long)__double_as_longlong(mi));
    }
}

// Reorders the entries of every (64-row slice, panel) run between row-major
// order and step-major ("jagged") order: step t holds entry t of each row of
// the slice that has more than t entries in the panel, rows in order.  One
// wavefront per run, lane l <-> row l: at step t the rows still active take
// consecutive positions, off_t + (active lanes below l).  Any number of
// matrices in one launch (grid.y = matrix); each run is written to the
// matrix's scratch arrays, four steps of loads in flight per lane, and
// k_relayout_back copies the scratch over the panel arrays.
__global__ void __launch_bounds__(kBlk) k_relayout_runs(const RelayoutJob* __restrict__ jobs, int to_jagged,
                                                       int64_t lo_e) {
    const RelayoutJob J = jobs[blockIdx.y];
    const int lane = threadIdx.x & 63;
    const int64_t n = J.n;
    const int64_t nslices = (n + 63) >> 6;
    const int64_t total = nslices * J.npan;
    const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t w = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); w < total; w += nw) {
        const int p = (int)(w / nslices);
        const int64_t slice = w - (int64_t)p * nslices;
        const uint32_t* rp = J.prow + (int64_t)p * (n + 1);
        const int64_t row0 = slice * 64, myrow = row0 + lane;
        const int64_t rend = row0 + 64 < n ? row0 + 64 : n;
        const uint32_t s = myrow < n ? rp[myrow] : rp[rend];
        const uint32_t e = myrow < n ? rp[myrow + 1] : rp[rend];
        const uint32_t len = e - s;
        const int64_t pb = J.pbase[p];
        const int64_t run = pb + rp[row0];
        if ((int64_t)rp[rend] - rp[row0] <= lo_e) continue;  // ordered by the greedy kernel
        int64_t off = 0;
        for (uint32_t t = 0;; t += 4) {
            uint64_t m[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) m[u] = __ballot(t + u < len);
            if (m[0] == 0) break;
            int64_t from[4], to[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int64_t jg = run + off + lanes_below(m[u]);
                const int64_t rm = pb + s + t + u;
                from[u] = to_jagged ? rm : jg;
                to[u] = to_jagged ? jg : rm;
                off += __popcll(m[u]);
            }
            uint16_t c[4] = {0, 0, 0, 0};
            double2 v[4] = {};
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (t + u < len) {
                    c[u] = J.col[from[u]];
                    v[u] = J.real ? make_double2(J.vr[from[u]], 0.0) : J.vc[from[u]];
                }
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (t + u < len) {
                    J.tcol[to[u]] = c[u];
                    if (J.real) J.tvr[to[u]] = v[u].x;
                    else J.tvc[to[u]] = v[u];
                }
        }
    }
}

// Greedy step order for the jagged layout (replaces the row order inside each
// step).  One wavefront per (slice, panel) run.  At step t every lane with
// more than t entries takes one of its entries; the lanes are served line by
// line: repeatedly the x cache line (kLineCols columns, 128 bytes of complex
// x) on which the most not-yet-served lanes still hold an entry is chosen
// (ties: the lowest line), and each of those lanes takes its lowest-column
// untaken entry on it.  Lanes gathering from one line at one step share one
// L1-to-L2 request.  The step positions are those of the plain jagged order
// (off_t + active lanes below l), so the SpMV kernels read either order; the
// order depends only on the run's (row, column) set, so converting back and
// forth reproduces it.  LDS per run (cap a power of two >= 64 entries):
// mask[cap] u64 (lanes holding an untaken entry per line), ecol[cap] u16
// (panel column per entry), dl[cap] u16 (sort buffer, then line index per
// entry), lines[cap] u16 (sorted distinct lines, then the live-line list),
// taken[cap] u8.  Runs of more than cap entries keep the plain order.
constexpr int kLineShift = 3;

__device__ void jag_copy_plain(const RelayoutJob& J, int64_t run, int64_t pb, uint32_t s, uint32_t len) {
    int64_t off = 0;
    for (uint32_t t = 0;; ++t) {
        const uint64_t m = __ballot(t < len);
        if (m == 0) break;
        if (t < len) {
            const int64_t to = run + off + lanes_below(m), from = pb + s + t;
            J.tcol[to] = J.col[from];
            if (J.real) J.tvr[to] = J.vr[from];
            else J.tvc[to] = J.vc[from];
        }
        off += __popcll(m);
    }
}

// This launch orders the runs with lo < E <= cap entries; runs above cap keep
// the plain order when plain_big is set and are left alone otherwise.
__global__ void __launch_bounds__(64) k_relayout_greedy(const RelayoutJob* __restrict__ jobs, int lo_e, int cap,
                                                        int plain_big, int* __restrict__ err) {
    extern __shared__ __align__(16) unsigned char lds_greedy[];
    uint64_t* mask = (uint64_t*)lds_greedy;
    uint16_t* ecol = (uint16_t*)(mask + cap);
    uint16_t* dl = ecol + cap;
    uint16_t* lines = dl + cap;
    uint8_t* taken = (uint8_t*)(lines + cap);
    const RelayoutJob J = jobs[blockIdx.y];
    const int lane = threadIdx.x;
    const int64_t n = J.n;
    const int64_t nslices = (n + 63) >> 6;
    const int64_t total = nslices * J.npan;
    for (int64_t w = blockIdx.x; w < total; w += gridDim.x) {
        const int p = (int)(w / nslices);
        const int64_t slice = w - (int64_t)p * nslices;
        const uint32_t* rp = J.prow + (int64_t)p * (n + 1);
        const int64_t row0 = slice * 64, myrow = row0 + lane;
        const int64_t rend = row0 + 64 < n ? row0 + 64 : n;
        const uint32_t r0 = rp[row0];
        const uint32_t s = myrow < n ? rp[myrow] : rp[rend];
        const uint32_t e = myrow < n ? rp[myrow + 1] : rp[rend];
        const uint32_t len = e - s;
        const int64_t pb = J.pbase[p];
        const int64_t run = pb + r0;
        const int64_t E = (int64_t)rp[rend] - r0;
        if (E <= lo_e) continue;
        if (E > cap) {
            if (plain_big) jag_copy_plain(J, run, pb, s, len);
            continue;
        }
        const int a = (int)(s - r0);
        uint32_t maxlen = len;
        for (int o = 32; o > 0; o >>= 1) {
            const uint32_t y = __shfl_xor(maxlen, o, 64);
            maxlen = y > maxlen ? y : maxlen;
        }
        int P2 = 64;
        while (P2 < E) P2 <<= 1;
        __syncthreads();  // the previous run's LDS reads are done
        for (int k = lane; k < P2; k += 64) {
            if (k < E) {
                const uint16_t c = J.col[run + k];
                ecol[k] = c;
                dl[k] = (uint16_t)(c >> kLineShift);
                taken[k] = 0;
            } else {
                dl[k] = 0xFFFF;
            }
        }
        __syncthreads();
        // bitonic sort of dl[0, P2)
        for (int size = 2; size <= P2; size <<= 1)
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                for (int i = lane; i < P2 / 2; i += 64) {
                    const int lo = 2 * i - (i & (stride - 1)), hi = lo + stride;
                    const bool up = (lo & size) == 0;
                    const uint16_t x = dl[lo], y = dl[hi];
                    if ((x > y) == up) {
                        dl[lo] = y;
                        dl[hi] = x;
                    }
                }
                __syncthreads();
            }
        // distinct lines
        int D = 0;
        for (int k0 = 0; k0 < E; k0 += 64) {
            const int k = k0 + lane;
            const bool first = k < E && (k == 0 || dl[k] != dl[k - 1]);
            const uint64_t b = __ballot(first);
            if (first) lines[D + lanes_below(b)] = dl[k];
            D += __popcll(b);
        }
        __syncthreads();
        for (int d = lane; d < D; d += 64) mask[d] = 0;
        __syncthreads();
        // line index of each of the lane's entries, lane bit on the line
        for (int k = a; k < a + (int)len; ++k) {
            const uint16_t ln = (uint16_t)(ecol[k] >> kLineShift);
            int lo = 0, hi = D - 1;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (lines[mid] < ln) lo = mid + 1;
                else hi = mid;
            }
            dl[k] = (uint16_t)lo;
            atomicOr((unsigned long long*)&mask[lo], 1ull << lane);
        }
        __syncthreads();
        // lines[] becomes the list of live line indices
        for (int d = lane; d < D; d += 64) lines[d] = (uint16_t)d;
        int live = D;
        __syncthreads();
        int64_t off = 0;
        for (uint32_t t = 0; t < maxlen; ++t) {
            const uint64_t act = __ballot(t < len);
            uint64_t U = act;
            int mine = -1;
            while (U) {
                int key = -1;
                for (int i = lane; i < live; i += 64) {
                    const int d = lines[i];
                    const int pc = __popcll(mask[d] & U);
                    const int kk = pc ? (pc << 16) | (0xFFFF - d) : -1;
                    key = kk > key ? kk : key;
                }
                for (int o = 32; o > 0; o >>= 1) {
                    const int y = __shfl_xor(key, o, 64);
                    key = y > key ? y : key;
                }
                if (key < 0) {  // no line serves an active lane: cannot happen
                    if (lane == 0) atomicOr(err, 1);
                    break;
                }
                const int bd = 0xFFFF - (key & 0xFFFF);
                const uint64_t take = mask[bd] & U;
                __syncthreads();  // every lane has read mask[bd]
                if ((take >> lane) & 1) {
                    int kb = -1, cnt = 0;
                    for (int k = a; k < a + (int)len; ++k)
                        if (!taken[k] && dl[k] == bd) {
                            ++cnt;
                            if (kb < 0 || ecol[k] < ecol[kb]) kb = k;
                        }
                    if (kb < 0) {
                        atomicOr(err, 2);
                    } else {
                        taken[kb] = 1;
                        mine = kb;
                        if (cnt == 1) atomicAnd((unsigned long long*)&mask[bd], ~(1ull << lane));
                    }
                }
                U &= ~take;
                __syncthreads();
            }
            if (t < len && mine >= 0) {
                const int64_t to = run + off + lanes_below(act), from = run + mine;
                J.tcol[to] = J.col[from];
                if (J.real) J.tvr[to] = J.vr[from];
                else J.tvc[to] = J.vc[from];
            }
            off += __popcll(act);
            // drop exhausted lines from the live list (in place, order kept)
            int nl = 0;
            for (int i0 = 0; i0 < live; i0 += 64) {
                const int i = i0 + lane;
                const int d = i < live ? lines[i] : 0;
                const bool keep = i < live && mask[d] != 0;
                const uint64_t b = __ballot(keep);
                __syncthreads();
                if (keep) lines[nl + lanes_below(b)] = (uint16_t)d;
                nl += __popcll(b);
                __syncthreads();
            }
            live = nl;
        }
    }
}

// Step-major order with R rows per lane (SpMV kind 5): the run of a
// 64R-row super-slice in a panel (its rows' entries, contiguous in the
// panel) is stored so that lane l walks rows l, l+64, .., l+64(R-1) one after
// the other as a single stream, and step t holds entry t of every lane's
// stream (active lanes in order).  Summing R rows per lane evens the lanes'
// stream lengths, so fewer lane-slots per step stay idle.
template <int R>
__global__ void __launch_bounds__(kBlk) k_relayout_rows(const RelayoutJob* __restrict__ jobs) {
    const RelayoutJob J = jobs[blockIdx.y];
    const int lane = threadIdx.x & 63;
    const int64_t n = J.n;
    const int64_t nss = (n + 64 * R - 1) / (64 * R);
    const int64_t total = nss * J.npan;
    const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t w = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); w < total; w += nw) {
        const int p = (int)(w / nss);
        const int64_t ss = w - (int64_t)p * nss;
        const uint32_t* rp = J.prow + (int64_t)p * (n + 1);
        const int64_t row0 = ss * 64 * R;
        const int64_t pb = J.pbase[p];
        const int64_t run = pb + rp[row0];
        uint32_t st[R], ln[R];
        uint32_t L = 0;
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int64_t row = row0 + lane + 64 * q;
            st[q] = row < n ? rp[row] : 0;
            ln[q] = row < n ? rp[row + 1] - st[q] : 0;
            L += ln[q];
        }
        int cur = 0;
        uint32_t k = 0;
        int64_t off = 0;
        for (uint32_t t = 0;; ++t) {
            const uint64_t m = __ballot(t < L);
            if (m == 0) break;
            if (t < L) {
                uint32_t lc = 0, sc = 0;
                bool ok = true;
                for (;;) {  // the lane's current row (rows without entries skipped; t < L keeps cur < R)
#pragma unroll
                    for (int q = 0; q < R; ++q)
                        if (q == cur) {
                            lc = ln[q];
                            sc = st[q];
                        }
                    if (k < lc) break;
                    ++cur;
                    k = 0;
                    if (cur >= R) {
                        ok = false;
                        break;
                    }
                }
                if (ok) {
                    const int64_t to = run + off + lanes_below(m), from = pb + sc + k;
                    J.tcol[to] = J.col[from];
                    if (J.real) J.tvr[to] = J.vr[from];
                    else J.tvc[to] = J.vc[from];
                    ++k;
                }
            }
            off += __popcll(m);
        }
    }
}

// Largest (slice, panel) run of every job: sizes the greedy order's LDS.
__global__ void __launch_bounds__(kBlk) k_run_max(const RelayoutJob* __restrict__ jobs, unsigned* __restrict__ out) {
    const RelayoutJob J = jobs[blockIdx.y];
    const int64_t n = J.n;
    const int64_t nslices = (n + 63) >> 6;
    const int64_t total = nslices * J.npan;
    unsigned m = 0;
    for (int64_t w = (int64_t)blockIdx.x * kBlk + threadIdx.x; w < total; w += (int64_t)gridDim.x * kBlk) {
        const int p = (int)(w / nslices);
        const int64_t slice = w - (int64_t)p * nslices;
        const uint32_t* rp = J.prow + (int64_t)p * (n + 1);
        const int64_t row0 = slice * 64, rend = row0 + 64 < n ? row0 + 64 : n;
        const unsigned E = rp[rend] - rp[row0];
        m = E > m ? E : m;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned y = __shfl_xor(m, o, 64);
        m = y > m ? y : m;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

__global__ void __launch_bounds__(kBlk) k_relayout_back(const RelayoutJob* __restrict__ jobs) {
    const RelayoutJob J = jobs[blockIdx.y];
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < J.nnz; k += (int64_t)gridDim.x * blockDim.x) {
        J.col[k] = J.tcol[k];
        if (J.real) J.vr[k] = J.tvr[k];
        else J.vc[k] = J.tvc[k];
    }
}

}  // namespace

namespace blqh {

namespace {
constexpr size_t kAlign = 256;

size_t greedy_lds_bytes(int cap) { return (size_t)cap * (8 + 2 + 2 + 2 + 1); }

// Most entries per run that the greedy jagged order can hold in LDS: 8192
// (120 KiB of LDS per workgroup) where the device allows it, else 4096 (60 KiB).
int greedy_cap_max() {
    static const int cap = [] {  // once, thread-safe
        int dev = 0, maxlds = 0;
        BLQH_HIP(hipGetDevice(&dev));
        BLQH_HIP(hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        int c = 4096;
        if ((size_t)maxlds >= greedy_lds_bytes(8192) &&
            hipFuncSetAttribute((const void*)k_relayout_greedy, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)greedy_lds_bytes(8192)) == hipSuccess)
            c = 8192;
        (void)hipGetLastError();
        return c;
    }();
    return cap;
}

// Runs are ordered in size classes (kGreedyFirst, doubling to kGreedySmall
// entries, then the rest up to the cap), each launch with the LDS its class
// needs (3.75 KiB for 256 entries, 30 KiB for 2048).
constexpr int kGreedyFirst = 256, kGreedySmall = 2048;

// LDS capacity for the large runs of these jobs: the smallest power of two
// >= 64 holding their largest run, at most greedy_cap_max() (larger runs keep
// the plain order).
int greedy_cap(const RelayoutJob* dj, int njobs, int64_t max_runs, hipStream_t st, unsigned& max_e) {
    max_e = 0;
    const int cmax = greedy_cap_max();
    if (cmax == 0) return 0;
    DevBuf<unsigned> mx(1);
    BLQH_HIP(hipMemsetAsync(mx.get(), 0, sizeof(unsigned), st));
    const int g = (int)std::max<int64_t>(1, std::min<int64_t>(ceil_div(max_runs, kBlk), 1024));
    hipLaunchKernelGGL(k_run_max, dim3(g, (unsigned)njobs), dim3(kBlk), 0, st, dj, mx.get());
    unsigned h = 0;
    BLQH_HIP(hipMemcpyAsync(&h, mx.get(), sizeof(unsigned), hipMemcpyDeviceToHost, st));
    BLQH_HIP(hipStreamSynchronize(st));
    max_e = h;
    int cap = 64;
    while (cap < cmax && (unsigned)cap < h) cap <<= 1;
    return cap;
}
size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }
size_t entry_bytes(const blqh_hamil* H) {
    const size_t m = (size_t)std::max<int64_t>(H->nnz, 1);
    return align_up(m * sizeof(uint16_t)) + align_up(m * (H->pan.real ? sizeof(double) : sizeof(double2)));
}
}  // namespace

size_t panel_layout_scratch_bytes(blqh_hamil* const* hs, int count) {
    size_t b = align_up((size_t)count * sizeof(RelayoutJob));
    for (int i = 0; i < count; ++i) b += entry_bytes(hs[i]);
    return b;
}

namespace {

// One conversion of every matrix in todo: row-major -> step-major (greedy or
// plain order, `rows` rows per lane) when jagged, else plain step-major with
// one row per lane -> row-major.
void relayout_pass(const std::vector<blqh_hamil*>& todo, bool jagged, bool greedy, int rows, void* scratch,
                   size_t scratch_bytes) {
    BLQH_REQUIRE(todo.size() <= 65535, BLQH_ERANGE, "too many matrices for one panel relayout");
    hipStream_t st = stream();
    const size_t need = panel_layout_scratch_bytes(todo.data(), (int)todo.size());
    DevBuf<char> own;
    char* sc = (char*)scratch;
    if (!sc || scratch_bytes < need) {
        own.alloc(need);
        sc = own.get();
    }
    std::vector<RelayoutJob> jobs(todo.size());
    size_t off = align_up(todo.size() * sizeof(RelayoutJob));
    int64_t max_runs = 1, max_nnz = 1;
    for (size_t i = 0; i < todo.size(); ++i) {
        blqh_hamil* H = todo[i];
        RelayoutJob& J = jobs[i];
        J.n = H->n;
        J.npan = H->pan.npan;
        J.real = H->pan.real ? 1 : 0;
        J.nnz = H->nnz;
        J.pbase = H->pan.pbase.get();
        J.prow = H->pan.prow.get();
        J.col = H->pan.pcol.get();
        J.vr = H->pan.pvr.get();
        J.vc = H->pan.pvc.get();
        const size_t m = (size_t)std::max<int64_t>(H->nnz, 1);
        J.tcol = (uint16_t*)(sc + off);
        off += align_up(m * sizeof(uint16_t));
        J.tvr = J.real ? (double*)(sc + off) : nullptr;
        J.tvc = J.real ? nullptr : (double2*)(sc + off);
        off += align_up(m * (J.real ? sizeof(double) : sizeof(double2)));
        max_runs = std::max<int64_t>(max_runs, ceil_div(H->n, 64) * H->pan.npan);
        max_nnz = std::max<int64_t>(max_nnz, H->nnz);
    }
    // earlier work on the library stream may still read this scratch; then a
    // synchronous copy of the job table (the host vector goes out of scope)
    BLQH_HIP(hipStreamSynchronize(st));
    copy_sync(sc, jobs.data(), jobs.size() * sizeof(RelayoutJob), hipMemcpyHostToDevice);
    const RelayoutJob* dj = (const RelayoutJob*)sc;
    const auto t0 = std::chrono::steady_clock::now();
    unsigned max_e = 0;
    const int cap = jagged && greedy && rows == 1 ? greedy_cap(dj, (int)jobs.size(), max_runs, st, max_e) : 0;
    DevBuf<int> err;
    const int gx = (int)std::max<int64_t>(1, std::min<int64_t>(ceil_div(max_runs, kBlk / 64), 4096));
    if (jagged && rows > 1) {
        const dim3 g(gx, (unsigned)jobs.size());
        if (rows == 2) hipLaunchKernelGGL(k_relayout_rows<2>, g, dim3(kBlk), 0, st, dj);
        else if (rows == 4) hipLaunchKernelGGL(k_relayout_rows<4>, g, dim3(kBlk), 0, st, dj);
        else if (rows == 8) hipLaunchKernelGGL(k_relayout_rows<8>, g, dim3(kBlk), 0, st, dj);
        else throw Error(BLQH_EINTERNAL, "unsupported rows per lane");
    } else if (cap > 0) {
        err.alloc(1);
        BLQH_HIP(hipMemsetAsync(err.get(), 0, sizeof(int), st));
        // one launch per size class of runs, each with the LDS of its largest
        // run: the kernel is bound by LDS latency and barriers, so smaller
        // classes run with more workgroups per CU
        const int gg = (int)std::max<int64_t>(1, std::min<int64_t>(max_runs, 8192));
        int lo = 0;
        for (int c = kGreedyFirst; lo < cap; c = c < kGreedySmall ? c * 2 : cap) {
            const int hi = std::min(c, cap);
            hipLaunchKernelGGL(k_relayout_greedy, dim3(gg, (unsigned)jobs.size()), dim3(64), greedy_lds_bytes(hi),
                               st, dj, lo, hi, 0, err.get());
            lo = hi;
        }
        if (max_e > (unsigned)cap)  // runs too large for LDS keep the plain order
            hipLaunchKernelGGL(k_relayout_runs, dim3(gx, (unsigned)jobs.size()), dim3(kBlk), 0, st, dj, 1,
                               (int64_t)cap);
    } else {
        hipLaunchKernelGGL(k_relayout_runs, dim3(gx, (unsigned)jobs.size()), dim3(kBlk), 0, st, dj, jagged ? 1 : 0,
                           (int64_t)-1);
    }
    const int gb = (int)std::max<int64_t>(1, std::min<int64_t>(ceil_div(max_nnz, kBlk), 4096));
    hipLaunchKernelGGL(k_relayout_back, dim3(gb, (unsigned)jobs.size()), dim3(kBlk), 0, st, dj);
    BLQH_HIP(hipGetLastError());
    if (cap > 0) {
        int herr = 0;
        BLQH_HIP(hipMemcpyAsync(&herr, err.get(), sizeof(int), hipMemcpyDeviceToHost, st));
        BLQH_HIP(hipStreamSynchronize(st));
        BLQH_REQUIRE(herr == 0, BLQH_EINTERNAL, "greedy jagged relayout left a lane without an entry");
        if (std::getenv("BLQH_TUNE_LOG"))
            std::fprintf(stderr, "blqh panel relayout: %zu matrices, greedy step order (cap %d): %.1f ms\n",
                         todo.size(), cap,
                         std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    }
    if (own.get()) BLQH_HIP(hipStreamSynchronize(st));  // the temporary is freed on return
    for (auto* H : todo) {
        H->pan.jagged = jagged;
        H->pan.greedy = jagged && greedy && rows == 1;
        H->pan.rows = jagged ? rows : 1;
    }
}

}  // namespace

void set_panel_layouts(blqh_hamil* const* hs, int count, bool jagged, void* scratch, size_t scratch_bytes,
                       bool greedy, int rows) {
    if (!jagged) rows = 1;
    if (rows > 1) greedy = false;
    // Every matrix ends in the requested order whatever happened to it before,
    // so a given caller always sums in the same order: the row-major order is
    // always the build's (ascending columns), the plain step order is derived
    // from it, and a greedy layout leaves step-major order by a fresh scatter
    // from the CSR.
    std::vector<blqh_hamil*> todo, back;
    hipStream_t st = stream();
    for (int i = 0; i < count; ++i) {
        blqh_hamil* H = hs[i];
        Panels& P = H->pan;
        if (!P.built || (P.jagged == jagged && (!jagged || (P.greedy == greedy && P.rows == rows)))) continue;
        if (P.jagged && (P.greedy || P.rows > 1)) {
            const int gr = (int)std::min<int64_t>(std::max<int64_t>(ceil_div(H->n, kBlk), 1), 65536);
            const CsrView A = spmv_csr(H);
            hipLaunchKernelGGL(k_panel_rescatter, dim3(gr), dim3(kBlk), 0, st, H->n, A.rowptr, A.col, A.val,
                               P.pbase.get(), P.prow.get(), P.real ? 1 : 0, P.pcol.get(),
                               P.pvr.get(), P.pvc.get());
            BLQH_HIP(hipGetLastError());
            P.jagged = P.greedy = false;
            P.rows = 1;
        } else if (P.jagged) {
            back.push_back(H);  // plain step order -> row-major by relayout
        }
        if (jagged) todo.push_back(H);
    }
    if (!back.empty()) relayout_pass(back, false, false, 1, scratch, scratch_bytes);
    if (!todo.empty()) relayout_pass(todo, jagged, greedy, rows, scratch, scratch_bytes);
}

// Step order of a single matrix's jagged panels: the plain order unless
// BLQH_JAGGED_ORDER=greedy (read at every layout change).  With the internal
// row order in place the greedy order no longer pays: measured on the 2.35M-row
// N_phi = 16 sector, its fused K1 is 3.06 ms against 2.93 ms in the plain
// order, and the greedy relayout costs 2.85 s once per matrix.
bool greedy_step_order() {
    const char* env = std::getenv("BLQH_JAGGED_ORDER");
    return env && std::strcmp(env, "greedy") == 0;
}

void set_panel_layout(blqh_hamil* H, bool jagged, void* scratch, size_t scratch_bytes, int rows) {
    set_panel_layouts(&H, 1, jagged, scratch, scratch_bytes, greedy_step_order(), rows);
}

size_t panel_layout_scratch_bytes(const blqh_hamil* H) {
    blqh_hamil* h = const_cast<blqh_hamil*>(H);
    return panel_layout_scratch_bytes(&h, 1);
}

// ---- internal order of the SpMV formats -----------------------------------------
//
// A basis state is a translation orbit: its members differ by one common
// rotation of every species word.  The SpMV formats order the rows by a key
// of the orbit that ignores which member the reference picked: rotate the
// state so that the lower layer's word x is smallest (ties: the upper
// layer's word y smallest), then key = (x, necklace class of y, position of
// y in its necklace).  Rows of one 64-row slice then share x, so the hops
// that move only lower-layer electrons send neighbouring rows to
// neighbouring columns (the upper-layer word only turns inside its
// necklace), and every other hop stays inside a few blocks of one x.  At
// 2.35M states (N_phi = 16, all densities) this raises the lanes that hold an
// entry at a step of the step-major panels from 24 to 54 of 64 and halves
// the distinct x cache lines a step gathers from.  The Lanczos recurrence is
// invariant under a permutation of the basis, so only the start vector and
// the returned states move between the orders.
const std::vector<int32_t>& internal_order(const blqh_basis* B) {
    std::lock_guard<std::mutex> lk(B->ord_mu);
    if (B->ord_done) return B->h_order;
    B->ord_done = true;
    const char* env = std::getenv("BLQH_ORDER");
    const int S = B->geo.S, nphi = B->geo.nphi;
    const int64_t n = B->n;
    if ((env && env[0] == '0') || B->sec.K < 0 || n < 2 || S < 2) return B->h_order;
    const int half = S / 2;  // species [0, half): upper layer; [half, S): lower layer
    const uint64_t m = nphi >= 64 ? ~0ull : ((1ull << nphi) - 1);
    auto group = [&](uint64_t w, int s0, int s1, int r) {
        uint64_t g = 0;
        for (int s = s0; s < s1; ++s) {
            const uint64_t x = (w >> (s * nphi)) & m;
            const uint64_t y = r ? (((x << r) | (x >> (nphi - r))) & m) : x;
            g |= y << ((s - s0) * nphi);
        }
        return g;
    };
    struct Key {
        uint64_t a, b;
        int32_t i;
    };
    std::vector<Key> keys(n);
    for (int64_t i = 0; i < n; ++i) {
        const uint64_t w = B->h_id[i];
        uint64_t bx = 0, by = 0, cls = 0;
        int rs = -1, r0 = 0;
        uint64_t ys[64];
        for (int r = 0; r < nphi; ++r) {
            const uint64_t x = group(w, half, S, r), y = group(w, 0, half, r);
            ys[r] = y;
            if (rs < 0 || x < bx || (x == bx && y < by)) { bx = x; by = y; rs = r; }
            if (r == 0 || y < cls) { cls = y; r0 = r; }
        }
        int ridx = 0;
        while (ys[(r0 + ridx) % nphi] != by) ++ridx;  // by is a rotation of cls: terminates
        keys[i] = Key{bx, (cls << 6) | (uint64_t)ridx, (int32_t)i};
    }
    std::sort(keys.begin(), keys.end(), [](const Key& p, const Key& q) {
        return p.a != q.a ? p.a < q.a : p.b != q.b ? p.b < q.b : p.i < q.i;
    });
    B->h_order.resize(n);
    for (int64_t r = 0; r < n; ++r) B->h_order[keys[r].i] = (int32_t)r;
    return B->h_order;
}

namespace {
constexpr int kOrdMaxRow = 4096;  // longest CSR row (the build's largest row hash)

__global__ void k_ord_len(int64_t n, const int32_t* __restrict__ iperm, const int64_t* __restrict__ rowptr,
                          int32_t* __restrict__ len) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t o = iperm[r];
        len[r] = (int32_t)(rowptr[o + 1] - rowptr[o]);
    }
}

// Row r of the internal CSR: reference row iperm[r] with its columns mapped
// through perm, placed in ascending order by rank (columns of a row are
// distinct).  One wavefront (workgroup) per row; rows longer than
// kOrdMaxRow are counted in *err and skipped.
__global__ void __launch_bounds__(64) k_ord_fill(int64_t n, const int32_t* __restrict__ iperm,
                                                 const int32_t* __restrict__ perm, const int64_t* __restrict__ rowptr,
                                                 const int32_t* __restrict__ col, const double2* __restrict__ val,
                                                 const int64_t* __restrict__ orp, int32_t* __restrict__ ocol,
                                                 double2* __restrict__ oval, int32_t* __restrict__ err) {
    __shared__ int32_t key[kOrdMaxRow];
    const int lane = threadIdx.x;
    for (int64_t r = blockIdx.x; r < n; r += gridDim.x) {
        const int64_t o = iperm[r];
        const int64_t s = rowptr[o];
        const int len = (int)(rowptr[o + 1] - s);
        if (len > kOrdMaxRow) {
            if (lane == 0) atomicAdd(err, 1);
            continue;  // uniform across the workgroup
        }
        __syncthreads();  // the previous row's keys are read
        for (int k = lane; k < len; k += 64) key[k] = perm[col[s + k]];
        __syncthreads();
        const int64_t d = orp[r];
        for (int k = lane; k < len; k += 64) {
            const int32_t c = key[k];
            int rank = 0;
            for (int q = 0; q < len; ++q) rank += key[q] < c ? 1 : 0;
            ocol[d + rank] = c;
            oval[d + rank] = val[s + k];
        }
    }
}

// internal[r] = reference[iperm[r]]
__global__ void k_ord_gather(int64_t n, const int32_t* __restrict__ iperm, const double2* __restrict__ src,
                             double2* __restrict__ dst) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
        dst[r] = src[iperm[r]];
}

// reference[iperm[r]] = internal[r]
__global__ void k_ord_scatter(int64_t n, const int32_t* __restrict__ iperm, const double2* __restrict__ src,
                              double2* __restrict__ dst) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
        dst[iperm[r]] = src[r];
}

int ord_grid(int64_t n) { return (int)std::min<int64_t>(std::max<int64_t>(ceil_div(n, kBlk), 1), 65536); }
}  // namespace

// The internal CSR of H from its CSR and the basis order (no-op for the identity order).
void build_order(blqh_hamil* H) {
    blqh_hamil::Order& O = H->ord;
    O.on = false;
    const std::vector<int32_t>& perm = internal_order(H->basis);
    if (perm.empty()) return;
    const int64_t n = H->n;
    BLQH_REQUIRE((int64_t)perm.size() == n, BLQH_EINTERNAL, "internal order and Hamiltonian sizes differ");
    hipStream_t st = stream();
    O.h_iperm.assign(n, 0);
    for (int64_t i = 0; i < n; ++i) O.h_iperm[perm[i]] = (int32_t)i;
    O.perm.alloc(n);
    O.iperm.alloc(n);
    BLQH_HIP(hipMemcpyAsync(O.perm.get(), perm.data(), n * 4, hipMemcpyHostToDevice, st));
    BLQH_HIP(hipMemcpyAsync(O.iperm.get(), O.h_iperm.data(), n * 4, hipMemcpyHostToDevice, st));
    DevBuf<int32_t> len(n);
    DevBuf<int32_t> err(1);
    BLQH_HIP(hipMemsetAsync(err.get(), 0, 4, st));
    hipLaunchKernelGGL(k_ord_len, dim3(ord_grid(n)), dim3(kBlk), 0, st, n, O.iperm.get(), H->rowptr.get(), len.get());
    BLQH_HIP(hipGetLastError());
    O.rowptr.alloc(n + 1);
    exclusive_scan(len.get(), O.rowptr.get(), n, st);
    O.col.alloc(std::max<int64_t>(H->nnz, 1));
    O.val.alloc(std::max<int64_t>(H->nnz, 1));
    const int g = (int)std::min<int64_t>(std::max<int64_t>(n, 1), (int64_t)runtime().num_cus * 40);
    hipLaunchKernelGGL(k_ord_fill, dim3(g), dim3(64), 0, st, n, O.iperm.get(), O.perm.get(), H->rowptr.get(),
                       H->col.get(), H->val.get(), O.rowptr.get(), O.col.get(), O.val.get(), err.get());
    BLQH_HIP(hipGetLastError());
    int32_t herr = 0;
    int64_t tot = 0;
    BLQH_HIP(hipMemcpyAsync(&herr, err.get(), 4, hipMemcpyDeviceToHost, st));
    BLQH_HIP(hipMemcpyAsync(&tot, O.rowptr.get() + n, 8, hipMemcpyDeviceToHost, st));
    BLQH_HIP(hipStreamSynchronize(st));
    BLQH_REQUIRE(herr == 0, BLQH_EINTERNAL, "a CSR row is longer than the internal-order kernel handles");
    BLQH_REQUIRE(tot == H->nnz, BLQH_EINTERNAL, "internal CSR size differs");
    O.on = true;
}

void order_to_internal(const blqh_hamil* H, const double2* src, double2* dst, hipStream_t st) {
    if (!H->ord.on) {
        if (src != dst) BLQH_HIP(hipMemcpyAsync(dst, src, H->n * sizeof(double2), hipMemcpyDeviceToDevice, st));
        return;
    }
    hipLaunchKernelGGL(k_ord_gather, dim3(ord_grid(H->n)), dim3(kBlk), 0, st, H->n, H->ord.iperm.get(), src, dst);
    BLQH_HIP(hipGetLastError());
}

void order_to_reference(const blqh_hamil* H, const double2* src, double2* dst, hipStream_t st) {
    if (!H->ord.on) {
        if (src != dst) BLQH_HIP(hipMemcpyAsync(dst, src, H->n * sizeof(double2), hipMemcpyDeviceToDevice, st));
        return;
    }
    hipLaunchKernelGGL(k_ord_scatter, dim3(ord_grid(H->n)), dim3(kBlk), 0, st, H->n, H->ord.iperm.get(), src, dst);
    BLQH_HIP(hipGetLastError());
}

void ensure_panels(blqh_hamil* H) {
    if (!H->pan.built) build_panels(H);
}

void build_panels(blqh_hamil* H) {
    build_order(H);
    const CsrView A = spmv_csr(H);
    hipStream_t st = stream();
    const int64_t n = H->n;
    const int npan = (int)std::max<int64_t>(1, ceil_div(n, (int64_t)1 << kPanelShift));
    Panels& P = H->pan;
    P.npan = npan;
    DevBuf<unsigned long long> mx(2);
    BLQH_HIP(hipMemsetAsync(mx.get(), 0, 16, st));
    const int g = (int)std::min<int64_t>(std::max<int64_t>(ceil_div(H->nnz, kBlk * 8), 1), 2048);
    hipLaunchKernelGGL(k_absmax, dim3(g), dim3(kBlk), 0, st, H->nnz, A.val, mx.get());
    unsigned long long hm[2];
    BLQH_HIP(hipMemcpyAsync(hm, mx.get(), 16, hipMemcpyDeviceToHost, st));
    BLQH_HIP(hipStreamSynchronize(st));
    double mre, mim;
    std::memcpy(&mre, &hm[0], 8);
    std::memcpy(&mim, &hm[1], 8);
    P.real = mim <= 1e-14 * std::max(mre, mim);
    DevBuf<int32_t> cnt((size_t)npan * n);
    DevBuf<int64_t> off((size_t)npan * n + 1);
    const int gr = (int)std::min<int64_t>(std::max<int64_t>(ceil_div(n, kBlk), 1), 65536);
    hipLaunchKernelGGL(k_panel_count, dim3(gr), dim3(kBlk), 0, st, n, npan, A.rowptr, A.col, cnt.get());
    BLQH_HIP(hipGetLastError());
    exclusive_scan(cnt.get(), off.get(), (int64_t)npan * n, st);
    P.prow.alloc((size_t)npan * (n + 1));
    P.pbase.alloc(npan + 1);
    const int gp = (int)std::min<int64_t>(ceil_div((int64_t)npan * (n + 1), kBlk), 65536);
    hipLaunchKernelGGL(k_panel_rows, dim3(gp), dim3(kBlk), 0, st, n, npan, off.get(), H->nnz, P.prow.get(),
                       P.pbase.get());
    BLQH_HIP(hipGetLastError());
    {
        // prow holds row offsets inside a panel as 32-bit values
        std::vector<int64_t> pb(npan + 1);
        BLQH_HIP(hipMemcpyAsync(pb.data(), P.pbase.get(), (npan + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, st));
        BLQH_HIP(hipStreamSynchronize(st));
        for (int p = 0; p < npan; ++p)
            if (pb[p + 1] - pb[p] > (int64_t)UINT32_MAX)
                throw Error(BLQH_ERANGE, "a column panel holds 2^32 or more entries (32-bit panel row offsets)");
    }
    P.pcol.alloc(std::max<int64_t>(H->nnz, 1));
    if (P.real) P.pvr.alloc(std::max<int64_t>(H->nnz, 1));
    else P.pvc.alloc(std::max<int64_t>(H->nnz, 1));
    hipLaunchKernelGGL(k_panel_scatter, dim3(gr), dim3(kBlk), 0, st, n, A.rowptr, A.col, A.val,
                       off.get(), P.real ? 1 : 0, P.pcol.get(), P.pvr.get(), P.pvc.get());
    BLQH_HIP(hipGetLastError());
    BLQH_HIP(hipStreamSynchronize(st));  // cnt / off are freed on return
    P.built = true;
}

}  // namespace blqh
