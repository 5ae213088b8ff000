// batch.hip -- many Lanczos ground-state solves advanced together.
//
// Ancestor: the reference's serial sweeps over sectors and parameter points
// (BLQH_without_spin/main.cpp:76-165, BLQH_with_spin/main.cpp:158-302), each
// of which runs lhamil's three passes (lanczos_hamiltonian.cpp:453-656) on one
// Hamiltonian at a time.  Here every problem of a batch (one resident
// Hamiltonian + one start vector) takes Lanczos step j in the same two
// launches:
//   K1  u = (H W_j)/beta_j - (beta_j/beta_{j-1}) W_{j-1},  alpha partials
//   K2  alpha_j = sum of the problem's K1 partials (fixed order),
//       W_{j+1} = u - alpha_j W_j / beta_j,               |W_{j+1}|^2 partials
//   K3  beta_{j+1} = sqrt(sum of the K2 partials)          (one wave per problem)
// The rows of every problem are cut into work items of kItemRows rows; a
// persistent grid walks the item list, so problems of any size share one
// launch and each item's partial sum has a fixed owner (bit-reproducible).
// Kernels read the step from a device counter and skip problems whose active
// flag is clear, so `check_every` steps are captured once in a hipGraph and
// replayed; between replays the host reads alpha/beta of every problem,
// runs the convergence test of the single-problem solver and clears the flags
// of the converged ones (data-dependent termination without a host round trip
// per step).  The Krylov vectors of every problem stay in HBM, so the ground
// state is one weighted sum per problem afterwards (no replay pass); the same
// pass forms the excited mix (second and third Ritz vectors), and both stay
// on the device in the reference row order as the start of a continuation
// solve at the next parameter point (k_bcont).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <condition_variable>
#include <cstring>
#include <exception>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>

#include "ritz.hpp"
#include "engine.hpp"
#include "mtstart.hpp"
#include "xcdplan.hpp"

namespace blqh {
void start_vector(int64_t n, uint32_t seed, std::vector<double2>& v);
double tridiag_lowest(int m, const double* a, const double* b, double* vec);
double tridiag_lowest_next(int m, const double* a, const double* b, double prev, double step);
}  // namespace blqh
void blqh_hamil_combine_impl(const blqh_hamil* base, const blqh_hamil* tun, double dsas, blqh_hamil* H);

namespace {

using namespace blqh;

constexpr int kItemRows = 256;  // rows per work item
constexpr int kBT = 256;        // threads per workgroup
constexpr int kSplit = 2;       // k_b1j: wavefronts per 64-row slice (measured: 2 beats 1 and 4)
constexpr int kBTj = kSplit * 256;  // threads per workgroup of k_b1j (an item's four slices)

// device descriptor of one problem
struct BProb {
    int64_t n;
    int32_t npan, real;
    const int64_t* pbase;
    const uint32_t* prow;
    const uint16_t* pcol;
    const double* pvr;
    const double2* pvc;
    int64_t item0;  // first work item
    int64_t voff;   // row offset in the pooled vectors
    double2* W;     // Krylov block: W_0 .. W_cap, n each
};

// per-step device state
struct BState {
    const BProb* probs;
    const int32_t* item_prob;  // work item -> problem
    int64_t nitems;
    const int32_t* xitems;     // k_b1j work lists: items grouped per XCD (real, complex, all members)
    const int32_t* xseg;       // [3][9] list offsets: kind (real, complex, all) x XCD
    const int32_t* active;     // [P]
    int32_t* jbase;            // step of the first kernel in a replay (advanced by k_bnext)
    int cap;                   // lambda
    double* alpha;             // [P][cap]
    double* beta;              // [P][cap + 1], beta[p][0] = 1
    double* invb;              // [P][cap + 1]
    double2* u;                // pooled rows
    double* part1;             // [nitems]
    double* part2;             // [nitems]
};

template <int NT = kBT>
__device__ __forceinline__ double block_sum(double v, double* sh) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0;
    for (int w = 0; w < NT / 64; ++w) t += sh[w];
    return t;
}

// sum of part[i0 .. i0+cnt) by one wavefront in a fixed order; every lane
// returns the same bits (the butterfly adds commutative pairs)
__device__ __forceinline__ double wave_sum(const double* part, int64_t i0, int64_t cnt, int lane) {
    double s = 0;
    for (int64_t i = lane; i < cnt; i += 64) s += part[i0 + i];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    return s;
}

// y_r = sum_k H_rk x_k for one row, LPR lanes (panel storage, any panel count)
template <int LPR>
__device__ __forceinline__ double2 prow_dot(const BProb& P, const double2* __restrict__ x, int64_t r, int lane) {
    double re = 0.0, im = 0.0;
    for (int pan = 0; pan < P.npan; ++pan) {
        const uint32_t* rp = P.prow + (int64_t)pan * (P.n + 1);
        const uint32_t s = rp[r], e = rp[r + 1];
        const int64_t base = P.pbase[pan];
        const double2* xp = x + (int64_t)pan * kPanelWidth;
        if (P.real) {
            for (uint32_t k = s + lane; k < e; k += LPR) {
                const int c = __builtin_nontemporal_load(&P.pcol[base + k]);
                const double v = __builtin_nontemporal_load(&P.pvr[base + k]);
                const double2 xv = xp[c];
                re += v * xv.x;
                im += v * xv.y;
            }
        } else {
            for (uint32_t k = s + lane; k < e; k += LPR) {
                const int c = __builtin_nontemporal_load(&P.pcol[base + k]);
                const double2 v = P.pvc[base + k];
                const double2 xv = xp[c];
                re += v.x * xv.x - v.y * xv.y;
                im += v.x * xv.y + v.y * xv.x;
            }
        }
    }
#pragma unroll
    for (int o = LPR / 2; o > 0; o >>= 1) {
        re += __shfl_xor(re, o, LPR);
        im += __shfl_xor(im, o, LPR);
    }
    return make_double2(re, im);
}

// K1 (LZ) or a plain product y = H x with partial Re<x|y> (!LZ: x and y are
// the pooled vectors xs / ys).
template <int LPR, bool LZ>
__global__ void __launch_bounds__(kBT) k_b1(BState S, int jrel, const double2* __restrict__ xs,
                                            double2* __restrict__ ys, double* __restrict__ part) {
    __shared__ double sh[kBT / 64];
    const int j = LZ ? *S.jbase + jrel : 0;
    if (LZ && j >= S.cap) return;
    constexpr int RPP = kBT / LPR;
    const int sub = threadIdx.x / LPR, lane = threadIdx.x % LPR;
    for (int64_t it = blockIdx.x; it < S.nitems; it += gridDim.x) {
        const int p = S.item_prob[it];
        if (!S.active[p]) continue;  // uniform over the workgroup
        const BProb P = S.probs[p];
        const int64_t r0 = (it - P.item0) * kItemRows;
        const int64_t r1 = r0 + kItemRows < P.n ? r0 + kItemRows : P.n;
        const double2* x;
        double2* y;
        const double2* Wp = nullptr;
        double ib = 1.0, cp = 0.0;
        if (LZ) {
            const int64_t ld = S.cap + 1;
            ib = S.invb[p * ld + j];
            cp = j > 0 ? S.beta[p * ld + j] * S.invb[p * ld + j - 1] : 0.0;
            x = P.W + (int64_t)j * P.n;
            Wp = j > 0 ? x - P.n : x;
            y = S.u + P.voff;
        } else {
            x = xs + P.voff;
            y = ys + P.voff;
        }
        double acc = 0.0;
        for (int64_t r = r0 + sub; r < r1; r += RPP) {
            const double2 h = prow_dot<LPR>(P, x, r, lane);
            if (lane == 0) {
                double ur = h.x, ui = h.y;
                if (LZ) {
                    ur = ib * h.x;
                    ui = ib * h.y;
                    if (j > 0) {
                        const double2 q = Wp[r];
                        ur -= cp * q.x;
                        ui -= cp * q.y;
                    }
                }
                y[r] = make_double2(ur, ui);
                const double2 w = x[r];
                acc += w.x * ur + w.y * ui;
            }
        }
        const double t = block_sum(acc, sh);
        if (threadIdx.x == 0) part[it] = t;
    }
}

// Same contract as k_b1 over panels in step-major order: kSplit wavefronts
// per 64-row slice of the item (waves kSplit*s .. kSplit*s + kSplit-1 own
// slice s, one lane per row), each summing every kSplit-th step group
// (jag_row_sum_part); the partial sums of waves 1.. go through LDS to wave 0
// of the slice, which adds them in wave order and runs the K1 epilogue.  KIND 0 / 1 walk only the members with real / complex panel
// values, KIND 2 walks all members and branches per work item (one launch per
// step for a mixed batch, so that the members of both kinds run side by side).
template <bool LZ, int KIND>
__global__ void __launch_bounds__(kBTj) k_b1j(BState S, int jrel, const double2* __restrict__ xs,
                                              double2* __restrict__ ys, double* __restrict__ part) {
    __shared__ double sh[kBTj / 64];
    __shared__ double2 red[kItemRows / 64][kSplit - 1][64];  // partial row sums of waves 1.. of each slice
    const int j = LZ ? *S.jbase + jrel : 0;
    if (LZ && j >= S.cap) return;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int sl = wv / kSplit, par = wv % kSplit;
    // Workgroups go to the XCDs round robin (workgroup b to XCD b mod 8, grid
    // a multiple of 8); each XCD walks its own list of whole problems, so the
    // x vectors it gathers from are those of the few problems its 32 CUs are
    // working on and stay in its L2.
    const int32_t* seg = S.xseg + KIND * 9 + (blockIdx.x & 7);
    const int32_t kend = seg[1];
    for (int32_t k = seg[0] + (int32_t)(blockIdx.x >> 3); k < kend; k += (int32_t)(gridDim.x >> 3)) {
        const int64_t it = S.xitems[k];
        const int p = S.item_prob[it];
        if (!S.active[p]) continue;  // uniform over the workgroup
        const BProb P = S.probs[p];
        const double2* x;
        double2* y;
        const double2* Wp = nullptr;
        double ib = 1.0, cp = 0.0;
        if (LZ) {
            const int64_t ld = S.cap + 1;
            ib = S.invb[p * ld + j];
            cp = j > 0 ? S.beta[p * ld + j] * S.invb[p * ld + j - 1] : 0.0;
            x = P.W + (int64_t)j * P.n;
            Wp = j > 0 ? x - P.n : x;
            y = S.u + P.voff;
        } else {
            x = xs + P.voff;
            y = ys + P.voff;
        }
        const int64_t n = P.n;
        const int64_t row0 = (it - P.item0) * kItemRows + sl * 64, myrow = row0 + lane;
        double ar = 0.0, ai = 0.0;
        if (row0 < n) {  // uniform over the wavefront
            const int64_t rend = row0 + 64 < n ? row0 + 64 : n;
            for (int pan = 0; pan < P.npan; ++pan) {
                const uint32_t* rp = P.prow + (int64_t)pan * (n + 1);
                const uint32_t s = myrow < n ? rp[myrow] : rp[rend];
                const uint32_t e = myrow < n ? rp[myrow + 1] : rp[rend];
                const int64_t run = P.pbase[pan] + (uint32_t)__builtin_amdgcn_readfirstlane((int)s);  // rp[row0]
                const double2* xp = x + (int64_t)pan * kPanelWidth;
                if (KIND == 2 ? P.real != 0 : KIND == 0)
                    jag_row_sum_part<true, kSplit>(P.pcol + run, P.pvr + run, P.pvc + run, xp, e - s, par, ar, ai);
                else
                    jag_row_sum_part<false, kSplit>(P.pcol + run, P.pvr + run, P.pvc + run, xp, e - s, par, ar, ai);
            }
        }
        if (par) red[sl][par - 1][lane] = make_double2(ar, ai);
        __syncthreads();
        double acc = 0.0;
        if (!par && myrow < n) {
#pragma unroll
            for (int q = 0; q < kSplit - 1; ++q) {
                const double2 o = red[sl][q][lane];
                ar += o.x;
                ai += o.y;
            }
            double ur = ar, ui = ai;
            if (LZ) {
                ur = ib * ar;
                ui = ib * ai;
                if (j > 0) {
                    const double2 q = Wp[myrow];
                    ur -= cp * q.x;
                    ui -= cp * q.y;
                }
            }
            y[myrow] = make_double2(ur, ui);
            const double2 w = x[myrow];
            acc = w.x * ur + w.y * ui;
        }
        const double t = block_sum<kBTj>(acc, sh);
        if (threadIdx.x == 0) part[it] = t;
    }
}

// ---- shared-scale groups ----------------------------------------------------
// G problems of one sector whose Hamiltonians are H0 + s_g T (the Delta_SAS
// points of one d) share one panel layout of H0 + T: the entries of T (those
// between words of different upper-layer occupation) carry bit 15 of their
// column (a sector of at most 32,768 states, one panel), and problem g scales
// them by s_g.  One pass over the entries serves all G problems: the column
// and value of an entry are loaded once and gathered against G vectors.
constexpr uint16_t kTunBit = 0x8000;

struct SGroups {
    const int32_t* gitem_grp;  // group work item -> group
    const int64_t* gitem0;     // [ngroups + 1] first group item
    const double* tscale;      // [P] scale of the flagged entries, problem p
};

// jag_row_sum_part for G vectors: lane sums of row `lane` of a 64-row slice
// over the steps of its wave (see jag_row_sum_part), entry values scaled by
// ts[g] where the column carries kTunBit.  act[g] == false skips vector g.
template <bool REAL, int NP, int G>
__device__ __forceinline__ void jag_row_sum_shared(const uint16_t* __restrict__ cr, const double* __restrict__ vr,
                                                   const double2* __restrict__ vc, const double2* const* xp,
                                                   const double* ts, const bool* act, uint32_t len, int par,
                                                   double* ar, double* ai) {
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
            const bool a = t + u < len;
            c[u] = a ? (int)__builtin_nontemporal_load(&cr[q]) : 0;
            if (REAL) {
                v[u] = make_double2(a ? __builtin_nontemporal_load(&vr[q]) : 0.0, 0.0);
            } else if (a) {
                const d2v_t w = __builtin_nontemporal_load(reinterpret_cast<const d2v_t*>(&vc[q]));
                v[u] = make_double2(w.x, w.y);
            } else {
                v[u] = make_double2(0.0, 0.0);
            }
        }
    };
    auto skip = [&](uint32_t t) {
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
        issue(t, c1, v1, any1);
        // one vector at a time: four gathers in flight per vector keep the
        // register count at 144 (all G x 4 at once: 162, and measured 9 % slower)
#pragma unroll
        for (int g = 0; g < G; ++g)
            if (act[g]) {
                double2 xv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) xv[u] = xp[g][c0[u] & (kTunBit - 1)];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const double sc = (c0[u] & kTunBit) ? ts[g] : 1.0;
                    const double wr = sc * v0[u].x;
                    if (REAL) {
                        ar[g] += wr * xv[u].x;
                        ai[g] += wr * xv[u].y;
                    } else {
                        const double wi = sc * v0[u].y;
                        ar[g] += wr * xv[u].x - wi * xv[u].y;
                        ai[g] += wr * xv[u].y + wi * xv[u].x;
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
}

// k_b1j for shared-scale groups: the work lists hold group items (256 rows
// of a sector); problems q G .. q G + G - 1 of group q share the sector's
// panels and own their vectors, scalars and partial sums (problem p's item
// li is item0 + li, as in k_b2 / k_b3).
template <bool LZ, int KIND, int G>
__global__ void __launch_bounds__(kBTj) k_b1s(BState S, SGroups Q, int jrel, const double2* __restrict__ xs,
                                              double2* __restrict__ ys, double* __restrict__ part) {
    __shared__ double sh[kBTj / 64];
    __shared__ double2 red[kItemRows / 64][kSplit - 1][G][64];
    const int j = LZ ? *S.jbase + jrel : 0;
    if (LZ && j >= S.cap) return;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int sl = wv / kSplit, par = wv % kSplit;
    const int32_t* seg = S.xseg + KIND * 9 + (blockIdx.x & 7);
    const int32_t kend = seg[1];
    for (int32_t k = seg[0] + (int32_t)(blockIdx.x >> 3); k < kend; k += (int32_t)(gridDim.x >> 3)) {
        const int64_t it = S.xitems[k];
        const int q = Q.gitem_grp[it];
        const int64_t li = it - Q.gitem0[q];
        bool act[G];
        bool anyact = false;
#pragma unroll
        for (int g = 0; g < G; ++g) {
            act[g] = S.active[q * G + g] != 0;
            anyact = anyact || act[g];
        }
        if (!anyact) continue;  // uniform over the workgroup
        const BProb P0 = S.probs[q * G];
        const int64_t n = P0.n;
        const double2* x[G];
        double2* y[G];
        const double2* Wp[G];
        double ib[G], cp[G], ts[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const BProb& Pg = S.probs[q * G + g];
            ts[g] = Q.tscale[q * G + g];
            ib[g] = 1.0;
            cp[g] = 0.0;
            if (LZ) {
                const int64_t ld = S.cap + 1;
                const int p = q * G + g;
                ib[g] = S.invb[p * ld + j];
                cp[g] = j > 0 ? S.beta[p * ld + j] * S.invb[p * ld + j - 1] : 0.0;
                x[g] = Pg.W + (int64_t)j * n;
                Wp[g] = j > 0 ? x[g] - n : x[g];
                y[g] = S.u + Pg.voff;
            } else {
                x[g] = xs + Pg.voff;
                Wp[g] = x[g];
                y[g] = ys + Pg.voff;
            }
        }
        const int64_t row0 = li * kItemRows + sl * 64, myrow = row0 + lane;
        double ar[G], ai[G];
#pragma unroll
        for (int g = 0; g < G; ++g) ar[g] = ai[g] = 0.0;
        if (row0 < n) {  // uniform over the wavefront; one panel (n <= 32768)
            const int64_t rend = row0 + 64 < n ? row0 + 64 : n;
            const uint32_t s = myrow < n ? P0.prow[myrow] : P0.prow[rend];
            const uint32_t e = myrow < n ? P0.prow[myrow + 1] : P0.prow[rend];
            const int64_t run = P0.pbase[0] + (uint32_t)__builtin_amdgcn_readfirstlane((int)s);
            if (KIND == 2 ? P0.real != 0 : KIND == 0)
                jag_row_sum_shared<true, kSplit, G>(P0.pcol + run, P0.pvr + run, P0.pvc + run, x, ts, act, e - s,
                                                    par, ar, ai);
            else
                jag_row_sum_shared<false, kSplit, G>(P0.pcol + run, P0.pvr + run, P0.pvc + run, x, ts, act, e - s,
                                                     par, ar, ai);
        }
        if (par) {
#pragma unroll
            for (int g = 0; g < G; ++g) red[sl][par - 1][g][lane] = make_double2(ar[g], ai[g]);
        }
        __syncthreads();
#pragma unroll
        for (int g = 0; g < G; ++g) {
            if (!act[g]) continue;  // uniform over the workgroup
            double acc = 0.0;
            if (!par && myrow < n) {
                double sr = ar[g], si = ai[g];
#pragma unroll
                for (int w = 0; w < kSplit - 1; ++w) {
                    const double2 o = red[sl][w][g][lane];
                    sr += o.x;
                    si += o.y;
                }
                double ur = sr, ui = si;
                if (LZ) {
                    ur = ib[g] * sr;
                    ui = ib[g] * si;
                    if (j > 0) {
                        const double2 qv = Wp[g][myrow];
                        ur -= cp[g] * qv.x;
                        ui -= cp[g] * qv.y;
                    }
                }
                y[g][myrow] = make_double2(ur, ui);
                const double2 w = x[g][myrow];
                acc = w.x * ur + w.y * ui;
            }
            const double t = block_sum<kBTj>(acc, sh);
            if (threadIdx.x == 0) part[S.probs[q * G + g].item0 + li] = t;
        }
    }
}

// T-entry flags of a one-panel sector in row-major panel order: bit 15 of
// the column of every entry between words of different upper-layer
// occupation (nu, internal row order); the flagged entries are counted.
__global__ void k_mark_tun(int64_t n, const int64_t* __restrict__ pbase, const uint32_t* __restrict__ prow,
                           uint16_t* __restrict__ pcol, const uint8_t* __restrict__ nu, int32_t* __restrict__ cnt) {
    int32_t c = 0;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = pbase[0];
        for (uint32_t k = prow[r]; k < prow[r + 1]; ++k) {
            const uint16_t col = pcol[b + k];
            if (nu[col] != nu[r]) {
                pcol[b + k] = (uint16_t)(col | kTunBit);
                ++c;
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(cnt, c);
}

// K2: alpha_j, W_{j+1} = u - alpha_j W_j / beta_j, |W_{j+1}|^2 partials
__global__ void __launch_bounds__(kBT) k_b2(BState S, int jrel) {
    __shared__ double sh[kBT / 64];
    __shared__ double sa;
    const int j = *S.jbase + jrel;
    if (j >= S.cap) return;
    const int64_t ld = S.cap + 1;
    for (int64_t it = blockIdx.x; it < S.nitems; it += gridDim.x) {
        const int p = S.item_prob[it];
        if (!S.active[p]) continue;
        const BProb P = S.probs[p];
        const int64_t nit = (P.n + kItemRows - 1) / kItemRows;
        const double ib = S.invb[p * ld + j];
        if (threadIdx.x < 64) {
            const double s = wave_sum(S.part1, P.item0, nit, threadIdx.x);
            if (threadIdx.x == 0) {
                sa = s * ib;
                if (it == P.item0) S.alpha[p * S.cap + j] = s * ib;
            }
        }
        __syncthreads();
        const double a = sa * ib;
        const int64_t r = (it - P.item0) * kItemRows + threadIdx.x;
        double acc = 0.0;
        if (r < P.n) {
            const double2 w = P.W[(int64_t)j * P.n + r], x = S.u[P.voff + r];
            const double2 v = make_double2(x.x - a * w.x, x.y - a * w.y);
            P.W[(int64_t)(j + 1) * P.n + r] = v;
            acc = v.x * v.x + v.y * v.y;
        }
        const double t = block_sum(acc, sh);
        if (threadIdx.x == 0) S.part2[it] = t;
        __syncthreads();  // sa is reused by the next item
    }
}

// K3: beta_{j+1}, 1/beta_{j+1}; one wavefront per problem
__global__ void __launch_bounds__(64) k_b3(BState S, int jrel, int nprob) {
    const int j = *S.jbase + jrel;
    if (j >= S.cap) return;
    const int p = blockIdx.x;
    if (p >= nprob || !S.active[p]) return;
    const BProb P = S.probs[p];
    const int64_t nit = (P.n + kItemRows - 1) / kItemRows;
    const double s = wave_sum(S.part2, P.item0, nit, threadIdx.x);
    if (threadIdx.x == 0) {
        const int64_t ld = S.cap + 1;
        const double b = sqrt(s);
        S.beta[p * ld + j + 1] = b;
        S.invb[p * ld + j + 1] = b > 0 ? 1.0 / b : 0.0;
    }
}

// end of a replay: the next replay starts `count` steps later
__global__ void k_bnext(int32_t* jbase, int count) {
    if (threadIdx.x == 0) *jbase += count;
}

// psi_r = sum_l coef_l conj(W_l[r]) over the problem's m vectors; |psi|^2 partials
// Seeded start vectors of a batch on the device: one workgroup per problem
// draws its std::mt19937 entries (mt_start_block) into gen; k_bstart_scale
// then writes dst[r] = gen[iperm[r]] / s (the member's internal row order;
// identity without one), s = sqrt of the problem's entry sum taken on the
// host as start_vector does: the same values as the host's start vectors.
struct BStart {
    int64_t n;
    uint32_t seed;
    double2* gen;
    double2* dst;
    const int32_t* iperm;
};

__global__ void __launch_bounds__(1024) k_bmt_start(const BStart* __restrict__ jobs, double* __restrict__ nrm2) {
    const BStart J = jobs[blockIdx.x];
    mt_start_block(J.n, J.seed, J.gen, nrm2 + blockIdx.x);
}

__global__ void __launch_bounds__(kBT) k_bstart_scale(const BStart* __restrict__ jobs, const double* __restrict__ s) {
    const BStart J = jobs[blockIdx.y];
    const double sj = s[blockIdx.y];
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < J.n; r += (int64_t)gridDim.x * blockDim.x) {
        double2 z = J.gen[J.iperm ? J.iperm[r] : r];
        z.x /= sj;
        z.y /= sj;
        J.dst[r] = z;
    }
}

// Continuation starts: the previous solve's ground state, kept in the
// reference row order (prev), conjugated (psi = sum_l c_l conj(v_l) is the
// eigenvector of conj(H), so conj(psi) is that of H), plus mix times the
// seeded unit vector: dst[r] = conj(prev[ip(r)]) + mix / |gen| * gen[ip(r)]
// + xmix * conj(prevx[ip(r)]) (ip = the member's internal order, the
// identity without one; prevx the previous solve's excited mix).  One
// 1024-thread workgroup per problem; nrm2 gets |dst|^2, the same fixed-order
// sum on every call.
struct BCont {
    int64_t n;
    const double2* prev;
    const double2* gen;  // seeded entries, reference order (read when mix > 0)
    const double2* prevx;  // the previous solve's excited mix, reference order (read when xmix > 0)
    double2* dst;
    const int32_t* iperm;
    double mix;          // start_mix / |gen|
    double xmix;
};

__global__ void __launch_bounds__(1024) k_bcont(const BCont* __restrict__ jobs, double* __restrict__ nrm2) {
    __shared__ double sh[1024 / 64];
    const BCont J = jobs[blockIdx.x];
    double acc = 0.0;
    for (int64_t r = threadIdx.x; r < J.n; r += 1024) {
        const int64_t q = J.iperm ? J.iperm[r] : r;
        const double2 a = J.prev[q];
        double2 z = make_double2(a.x, -a.y);
        if (J.mix > 0) {
            const double2 g = J.gen[q];
            z.x += J.mix * g.x;
            z.y += J.mix * g.y;
        }
        if (J.xmix > 0) {
            const double2 g = J.prevx[q];
            z.x += J.xmix * g.x;
            z.y -= J.xmix * g.y;
        }
        J.dst[r] = z;
        acc += z.x * z.x + z.y * z.y;
    }
    const double t = block_sum<1024>(acc, sh);
    if (threadIdx.x == 0) nrm2[blockIdx.x] = t;
}

// Ground states back to the reference row order for the copy to the host:
// dst[iperm[r]] = src[r] (a member in its internal order), dst[r] = src[r]
// otherwise; one job per problem (blockIdx.y).
struct BPsi {
    int64_t n;
    const int32_t* iperm;
    const double2* src;
    double2* dst;
};

__global__ void __launch_bounds__(kBT) k_bpsi_ref(const BPsi* __restrict__ jobs) {
    const BPsi J = jobs[blockIdx.y];
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < J.n; r += (int64_t)gridDim.x * blockDim.x)
        J.dst[J.iperm ? J.iperm[r] : r] = J.src[r];
}

__global__ void __launch_bounds__(kBT) k_bcombine(BState S, const double* __restrict__ coef,
                                                  const double* __restrict__ coefx, const int32_t* __restrict__ mvec,
                                                  double2* __restrict__ psi, double2* __restrict__ xs) {
    __shared__ double sh[kBT / 64];
    for (int64_t it = blockIdx.x; it < S.nitems; it += gridDim.x) {
        const int p = S.item_prob[it];
        const BProb P = S.probs[p];
        const int m = mvec[p];
        const double* cf = coef + (int64_t)p * S.cap;
        const double* cx = coefx + (int64_t)p * S.cap;
        const int64_t r = (it - P.item0) * kItemRows + threadIdx.x;
        double acc = 0.0;
        if (r < P.n) {
            double re = 0.0, im = 0.0, xr = 0.0, xi = 0.0;
            for (int l = 0; l < m; ++l) {
                const double2 w = P.W[(int64_t)l * P.n + r];
                re += cf[l] * w.x;
                im -= cf[l] * w.y;
                xr += cx[l] * w.x;
                xi -= cx[l] * w.y;
            }
            psi[P.voff + r] = make_double2(re, im);
            xs[P.voff + r] = make_double2(xr, xi);
            acc = re * re + im * im;
        }
        const double t = block_sum(acc, sh);
        if (threadIdx.x == 0) S.part1[it] = t;
    }
}

__global__ void __launch_bounds__(kBT) k_bscale(BState S, const double* __restrict__ scale, double2* __restrict__ psi) {
    for (int64_t it = blockIdx.x; it < S.nitems; it += gridDim.x) {
        const int p = S.item_prob[it];
        const BProb P = S.probs[p];
        const int64_t r = (it - P.item0) * kItemRows + threadIdx.x;
        if (r < P.n) {
            double2 z = psi[P.voff + r];
            z.x *= scale[p];
            z.y *= scale[p];
            psi[P.voff + r] = z;
        }
    }
}

}  // namespace

namespace {
// Pooled per-problem scalars and vectors of a batched solve.
// Host worker threads for the per-replay convergence tests, kept for the
// batch's lifetime: starting 16 threads costs about 0.5 ms, as much as a
// replay of two Lanczos steps.  run(nt, f) calls f(t) for t < nt, the caller
// taking t = 0, and returns when every call has returned.
class HostPool {
public:
    explicit HostPool(int workers) {
        for (int w = 1; w <= workers; ++w) th_.emplace_back([this, w] { loop(w); });
    }
    ~HostPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : th_) t.join();
    }
    int size() const { return (int)th_.size() + 1; }
    void run(int nt, const std::function<void(int)>& f) {
        nt = std::max(1, std::min(nt, size()));
        {
            std::lock_guard<std::mutex> lk(mu_);
            job_ = &f;
            nt_ = nt;
            left_ = nt - 1;
            ++gen_;
        }
        if (nt > 1) cv_.notify_all();
        f(0);
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [this] { return left_ == 0; });
        job_ = nullptr;
    }

private:
    void loop(int w) {
        int seen = 0;
        while (true) {
            const std::function<void(int)>* job;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
                if (w >= nt_) continue;
                job = job_;
            }
            (*job)(w);
            std::lock_guard<std::mutex> lk(mu_);
            if (--left_ == 0) done_.notify_one();
        }
    }
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_, done_;
    const std::function<void(int)>* job_ = nullptr;
    int nt_ = 0, left_ = 0, gen_ = 0;
    bool stop_ = false;
};

struct Pools {
    DevBuf<BProb> probs;
    DevBuf<int32_t> active, jbase, mvec;
    DevBuf<double> alpha, beta, invb, part1, part2, coef, coefx, scale;
    DevBuf<double2> u, psi, W;
    // the previous solve's excited mix (normalised sum of its second and
    // third lowest Ritz vectors): internal order, then reference order
    DevBuf<double2> xs, xref;
};
}  // namespace

struct blqh_batch {
    std::vector<blqh_hamil*> hs;
    std::vector<BProb> h_probs;
    std::vector<int64_t> item0;  // [P + 1]
    int64_t rows = 0, nitems = 0, nnz = 0;
    DevBuf<int32_t> item_prob;
    DevBuf<int32_t> xitems, xseg;    // per-XCD work lists of k_b1j (rebuilt by refresh_probs)
    std::vector<int32_t> h_xitems, h_xseg;
    int lpr = 0;                 // lanes per row of k_b1 (autotuned)
    bool has_real = false, has_complex = false;  // value kinds among the members
    int grid = 0;
    // kept between solves of the same lambda: device pools, the captured
    // replay (for one check interval and lane count) and pinned staging
    Pools pl;
    int pool_cap = -1;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int g_chk = -1, g_lpr = -1, g_cap = -1, g_kinds = -1;
    char* pin = nullptr;
    double2* psi_pin = nullptr;  // pinned staging of the ground states (b->rows entries, on first use)
    // pl.u holds every problem's ground state of the last finished solve in
    // the reference row order (the start of a continuation solve)
    bool cont_ok = false;
    DevBuf<char> relayout;       // scratch for reordering members' panels
    // Shared-scale groups (blqh_batch_create_scaled); G = 0: every member
    // has its own matrix.  Problem p = s G + g is H0[s] + scale[g] T[s];
    // hs[p] is the sector's template tpl[s] (owned, flagged T entries).
    int G = 0;
    std::vector<blqh_hamil*> tpl;
    std::vector<double> tscale;   // [P]
    std::vector<int64_t> gitem0;  // [nsector + 1] group work items
    int64_t ngitems = 0;
    DevBuf<int32_t> gitem_grp;
    DevBuf<int64_t> dgitem0;
    DevBuf<double> dtscale;
    std::unique_ptr<HostPool> hpool;  // convergence-test threads (batches of 32 or more problems)
    ~blqh_batch() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        if (pin) (void)hipHostFree(pin);
        if (psi_pin) (void)hipHostFree(psi_pin);
        for (blqh_hamil* H : tpl) delete H;
    }
};

namespace {

// Planner view of the members (all, or those with live[p] set); with
// shared-scale groups, of the groups (live when one of their problems is).
std::vector<XcdMember> xcd_members(const blqh_batch* b, const std::vector<char>* live = nullptr) {
    std::vector<XcdMember> mem;
    if (b->G) {
        const int ns = (int)b->tpl.size();
        mem.reserve(ns);
        for (int q = 0; q < ns; ++q) {
            bool on = !live;
            for (int g = 0; g < b->G && !on; ++g) on = (*live)[q * b->G + g] != 0;
            if (on)
                mem.push_back({b->gitem0[q], b->gitem0[q + 1], b->tpl[q]->nnz * b->G, 1, b->tpl[q]->pan.real ? 1 : 0});
        }
        return mem;
    }
    const int P = (int)b->hs.size();
    mem.reserve(P);
    for (int p = 0; p < P; ++p)
        if (!live || (*live)[p])
            mem.push_back({b->item0[p], b->item0[p + 1], b->hs[p]->nnz, b->hs[p]->pan.npan,
                           b->hs[p]->pan.real ? 1 : 0});
    return mem;
}

// Work lists of k_b1j (xcd_lists), uploaded when they change.
void build_xcd_lists(blqh_batch* b) {
    std::vector<int32_t> items, seg;
    xcd_lists(xcd_members(b), items, seg);
    if (items == b->h_xitems && seg == b->h_xseg) return;
    b->h_xitems.swap(items);
    b->h_xseg.swap(seg);
    // ordered after every kernel queued so far; the host vectors are not
    // touched again before the copies have completed
    hipStream_t st = stream();
    BLQH_HIP(hipMemcpyAsync(b->xitems.get(), b->h_xitems.data(), b->h_xitems.size() * 4, hipMemcpyHostToDevice, st));
    BLQH_HIP(hipMemcpyAsync(b->xseg.get(), b->h_xseg.data(), 27 * 4, hipMemcpyHostToDevice, st));
    BLQH_HIP(hipStreamSynchronize(st));
}

// Row-major panel layout and current panel buffers of every member.
void refresh_probs(blqh_batch* b) {
    b->has_real = b->has_complex = false;
    for (blqh_hamil* H : b->hs) (H->pan.real ? b->has_real : b->has_complex) = true;
    // k_b1j reads step-major runs, k_b1 row-major ones: one relayout launch for all members
    // plain step order: a rebind re-lays out every member at each parameter point
    if (b->G)  // the templates, once each
        set_panel_layouts(b->tpl.data(), (int)b->tpl.size(), true, b->relayout.get(), b->relayout.bytes(), false);
    else
        set_panel_layouts(b->hs.data(), (int)b->hs.size(), b->lpr == 64, b->relayout.get(), b->relayout.bytes(),
                          false);
    build_xcd_lists(b);
    for (size_t p = 0; p < b->hs.size(); ++p) {
        blqh_hamil* H = b->hs[p];
        BProb& d = b->h_probs[p];
        d.pbase = H->pan.pbase.get();
        d.prow = H->pan.prow.get();
        d.pcol = H->pan.pcol.get();
        d.pvr = H->pan.pvr.get();
        d.pvc = H->pan.pvc.get();
    }
}

SGroups make_groups(const blqh_batch* b) {
    return SGroups{b->gitem_grp.get(), b->dgitem0.get(), b->dtscale.get()};
}

template <bool LZ, int KIND>
void launch_b1s_kind(const blqh_batch* b, const BState& S, int jrel, const double2* xs, double2* ys, double* part,
                     hipStream_t st) {
    const dim3 g(b->grid), tj(kBTj);
    const SGroups Q = make_groups(b);
    switch (b->G) {
        case 1: hipLaunchKernelGGL((k_b1s<LZ, KIND, 1>), g, tj, 0, st, S, Q, jrel, xs, ys, part); break;
        case 2: hipLaunchKernelGGL((k_b1s<LZ, KIND, 2>), g, tj, 0, st, S, Q, jrel, xs, ys, part); break;
        case 3: hipLaunchKernelGGL((k_b1s<LZ, KIND, 3>), g, tj, 0, st, S, Q, jrel, xs, ys, part); break;
        case 4: hipLaunchKernelGGL((k_b1s<LZ, KIND, 4>), g, tj, 0, st, S, Q, jrel, xs, ys, part); break;
        default: throw Error(BLQH_EINTERNAL, "unsupported group size");
    }
}

template <bool LZ>
void launch_b1(const blqh_batch* b, const BState& S, int jrel, const double2* xs, double2* ys, double* part,
               hipStream_t st) {
    const dim3 g(b->grid), t(kBT), tj(kBTj);
    if (b->G) {  // shared-scale groups: one pass over each sector's entries for all its problems
        if (b->has_real && b->has_complex) {
            launch_b1s_kind<LZ, 2>(b, S, jrel, xs, ys, part, st);
        } else {
            if (b->has_real) launch_b1s_kind<LZ, 0>(b, S, jrel, xs, ys, part, st);
            if (b->has_complex) launch_b1s_kind<LZ, 1>(b, S, jrel, xs, ys, part, st);
        }
        return;
    }
    switch (b->lpr) {
        case 64:
            if (b->has_real && b->has_complex) {  // measured faster than one launch per kind
                hipLaunchKernelGGL((k_b1j<LZ, 2>), g, tj, 0, st, S, jrel, xs, ys, part);
                break;
            }
            if (b->has_real) hipLaunchKernelGGL((k_b1j<LZ, 0>), g, tj, 0, st, S, jrel, xs, ys, part);
            if (b->has_complex) hipLaunchKernelGGL((k_b1j<LZ, 1>), g, tj, 0, st, S, jrel, xs, ys, part);
            break;
        case 4: hipLaunchKernelGGL((k_b1<4, LZ>), g, t, 0, st, S, jrel, xs, ys, part); break;
        case 8: hipLaunchKernelGGL((k_b1<8, LZ>), g, t, 0, st, S, jrel, xs, ys, part); break;
        case 16: hipLaunchKernelGGL((k_b1<16, LZ>), g, t, 0, st, S, jrel, xs, ys, part); break;
        case 32: hipLaunchKernelGGL((k_b1<32, LZ>), g, t, 0, st, S, jrel, xs, ys, part); break;
        default: throw Error(BLQH_EINTERNAL, "unsupported batched SpMV lanes");
    }
}

BState make_state(const blqh_batch* b, Pools& pl, int cap) {
    BState S{};
    S.probs = pl.probs.get();
    S.item_prob = b->item_prob.get();
    S.nitems = b->nitems;
    S.xitems = b->xitems.get();
    S.xseg = b->xseg.get();
    S.active = pl.active.get();
    S.jbase = pl.jbase.get();
    S.cap = cap;
    S.alpha = pl.alpha.get();
    S.beta = pl.beta.get();
    S.invb = pl.invb.get();
    S.u = pl.u.get();
    S.part1 = pl.part1.get();
    S.part2 = pl.part2.get();
    return S;
}

// time the plain batched product for each lane count and keep the fastest
void tune_batch(blqh_batch* b) {
    hipStream_t st = stream();
    Pools pl;
    const int P = (int)b->hs.size();
    pl.probs.alloc(P);
    pl.active.alloc(P);
    BLQH_HIP(hipMemcpyAsync(pl.probs.get(), b->h_probs.data(), P * sizeof(BProb), hipMemcpyHostToDevice, st));
    std::vector<int32_t> on(P, 1);
    BLQH_HIP(hipMemcpyAsync(pl.active.get(), on.data(), P * 4, hipMemcpyHostToDevice, st));
    DevBuf<double2> x(b->rows), y(b->rows);
    DevBuf<double> part(b->nitems);
    BLQH_HIP(hipMemsetAsync(x.get(), 0, b->rows * sizeof(double2), st));
    BState S = make_state(b, pl, 1);
    hipEvent_t e0, e1;
    BLQH_HIP(hipEventCreate(&e0));
    BLQH_HIP(hipEventCreate(&e1));
    float best = 1e30f;
    int bl = 8;
    for (int l : {4, 8, 16, 32, 64}) {  // 64: one lane per row, step-major panels (last: one relayout)
        b->lpr = l;
        refresh_probs(b);
        BLQH_HIP(hipMemcpyAsync(pl.probs.get(), b->h_probs.data(), P * sizeof(BProb), hipMemcpyHostToDevice, st));
        launch_b1<false>(b, S, 0, x.get(), y.get(), part.get(), st);
        BLQH_HIP(hipEventRecord(e0, st));
        for (int r = 0; r < 3; ++r) launch_b1<false>(b, S, 0, x.get(), y.get(), part.get(), st);
        BLQH_HIP(hipEventRecord(e1, st));
        BLQH_HIP(hipEventSynchronize(e1));
        float ms = 0;
        BLQH_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) { best = ms; bl = l; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->lpr = bl;
    refresh_probs(b);
}

}  // namespace

blqh_batch* blqh_batch_create_impl(blqh_hamil* const* hs, int count, int lanes) {
    BLQH_REQUIRE(lanes == 0 || lanes == 4 || lanes == 8 || lanes == 16 || lanes == 32 || lanes == 64, BLQH_EINVAL,
                 "unsupported lanes per row");
    BLQH_REQUIRE(count >= 1, BLQH_EINVAL, "empty batch");
    auto* b = new blqh_batch;
    try {
        b->hs.assign(hs, hs + count);
        b->item0.assign(count + 1, 0);
        std::vector<int32_t> ip;
        for (int p = 0; p < count; ++p) {
            blqh_hamil* H = hs[p];
            BLQH_REQUIRE(H != nullptr, BLQH_EINVAL, "null hamiltonian in batch");
            ensure_panels(H);
            BLQH_REQUIRE(H->pan.built && H->n > 0, BLQH_EINVAL, "batched solve needs a non-empty panel layout");
            BProb d{};
            d.n = H->n;
            d.npan = H->pan.npan;
            d.real = H->pan.real ? 1 : 0;
            d.item0 = b->nitems;
            d.voff = b->rows;
            const int64_t nit = ceil_div(H->n, kItemRows);
            for (int64_t i = 0; i < nit; ++i) ip.push_back(p);
            b->nitems += nit;
            b->rows += H->n;
            b->nnz += H->nnz;
            b->item0[p + 1] = b->nitems;
            b->h_probs.push_back(d);
        }
        BLQH_REQUIRE(b->nitems < (int64_t)1 << 31, BLQH_ERANGE, "batch too large");
        b->item_prob.alloc(ip.size());
        copy_sync(b->item_prob.get(), ip.data(), ip.size() * 4, hipMemcpyHostToDevice);
        // a multiple of 8 (one share per XCD in k_b1j)
        b->grid = (int)(8 * ceil_div(std::min<int64_t>(b->nitems, (int64_t)runtime().num_cus * 8), 8));
        b->xitems.alloc(2 * b->nitems);
        b->xseg.alloc(27);
        b->relayout.alloc(panel_layout_scratch_bytes(b->hs.data(), count));
        if (lanes) {
            b->lpr = lanes;
            refresh_probs(b);
        } else {
            refresh_probs(b);
            tune_batch(b);
        }
    } catch (...) {
        delete b;
        throw;
    }
    return b;
}

// New members of the same shapes (a sweep's next parameter point): the work
// items, pools, captured replay and lane choice are kept.
void blqh_batch_rebind_impl(blqh_batch* b, blqh_hamil* const* hs, int count) {
    BLQH_REQUIRE(b->G == 0, BLQH_EINVAL, "a shared-scale batch is rebound with blqh_batch_rebind_scaled");
    BLQH_REQUIRE(count == (int)b->hs.size(), BLQH_EINVAL, "rebind needs as many Hamiltonians as the batch has");
    int64_t nnz = 0;
    for (int p = 0; p < count; ++p) {
        BLQH_REQUIRE(hs[p] != nullptr, BLQH_EINVAL, "null hamiltonian in batch");
        ensure_panels(hs[p]);
        BLQH_REQUIRE(hs[p]->n == b->h_probs[p].n && hs[p]->pan.npan == b->h_probs[p].npan, BLQH_EINVAL,
                     "rebind needs Hamiltonians of the same sizes, in the same order");
        nnz += hs[p]->nnz;
    }
    b->nnz = nnz;
    b->hs.assign(hs, hs + count);
    for (int p = 0; p < count; ++p) b->h_probs[p].real = hs[p]->pan.real ? 1 : 0;
    const size_t scratch = panel_layout_scratch_bytes(b->hs.data(), count);
    if (scratch > b->relayout.bytes()) b->relayout.alloc(scratch);
    refresh_probs(b);
}

namespace {
// H0 + 1 T of one sector (blqh_hamil_combine) with bit 15 set on the column
// of every T entry, in row-major panel order.  The tunnelling terms are the
// only ones that move an electron between the layers, so T's entries are
// exactly those between words of different upper-layer occupation; their
// count must equal T's.
blqh_hamil* make_template(const blqh_hamil* base, const blqh_hamil* tun) {
    BLQH_REQUIRE(base != nullptr && tun != nullptr, BLQH_EINVAL, "null hamiltonian in batch");
    BLQH_REQUIRE(base->basis->geo.S >= 2, BLQH_EINVAL, "a shared-scale batch needs bilayer sectors");
    BLQH_REQUIRE(base->n <= (int64_t)kTunBit, BLQH_EINVAL,
                 "a shared-scale batch takes sectors of at most 32768 states");
    auto* H = new blqh_hamil;
    try {
        blqh_hamil_combine_impl(base, tun, 1.0, H);
        ensure_panels(H);
        BLQH_REQUIRE(H->pan.built && H->pan.npan == 1, BLQH_EINTERNAL, "template is not a one-panel matrix");
        set_panel_layout(H, false);  // row-major: the marking kernel walks rows
        const blqh_basis* B = H->basis;
        const int64_t n = H->n;
        const uint64_t upper = lowbits(B->geo.S * B->geo.nphi / 2);
        std::vector<uint8_t> nu(n);
        for (int64_t r = 0; r < n; ++r) {
            const int64_t i = H->ord.on ? H->ord.h_iperm[r] : r;
            nu[r] = (uint8_t)popc(B->h_id[i] & upper);
        }
        hipStream_t st = stream();
        DevBuf<uint8_t> dnu(n);
        DevBuf<int32_t> cnt(1);
        BLQH_HIP(hipMemcpyAsync(dnu.get(), nu.data(), n, hipMemcpyHostToDevice, st));
        BLQH_HIP(hipMemsetAsync(cnt.get(), 0, 4, st));
        const int g = (int)std::min<int64_t>(std::max<int64_t>(ceil_div(n, 256), 1), 4096);
        hipLaunchKernelGGL(k_mark_tun, dim3(g), dim3(256), 0, st, n, H->pan.pbase.get(), H->pan.prow.get(),
                           H->pan.pcol.get(), dnu.get(), cnt.get());
        BLQH_HIP(hipGetLastError());
        int32_t hc = 0;
        copy_sync(&hc, cnt.get(), 4, hipMemcpyDeviceToHost);
        BLQH_REQUIRE(hc == tun->nnz, BLQH_EINTERNAL, "the tunnelling entries are not the layer-changing entries");
    } catch (...) {
        delete H;
        throw;
    }
    return H;
}

// Templates of nsec sectors, built on up to 8 host threads: one sector's
// merge, panels and marking are small launches with host waits between
// them, so several sectors on their own streams keep the GPU busy.  On an
// error every built template is freed and the first error is raised.
std::vector<blqh_hamil*> make_templates(const blqh_hamil* const* base, const blqh_hamil* const* tun, int nsec) {
    std::vector<blqh_hamil*> out(nsec, nullptr);
    const int nt = std::max(1, std::min(nsec, 8));
    std::vector<std::exception_ptr> err(nt);
    auto work = [&](int t) {
        try {
            for (int q = t; q < nsec; q += nt) out[q] = make_template(base[q], tun[q]);
        } catch (...) {
            err[t] = std::current_exception();
        }
    };
    if (nt == 1) {
        work(0);
    } else {
        std::vector<std::thread> th;
        for (int t = 1; t < nt; ++t) th.emplace_back(work, t);
        work(0);
        for (auto& x : th) x.join();
    }
    for (auto& e : err)
        if (e) {
            for (blqh_hamil* H : out) delete H;
            std::rethrow_exception(e);
        }
    return out;
}

}  // namespace

blqh_batch* blqh_batch_create_scaled_impl(const blqh_hamil* const* base, const blqh_hamil* const* tun, int nsec,
                                          const double* scales, int nscale) {
    BLQH_REQUIRE(nsec >= 1, BLQH_EINVAL, "empty batch");
    BLQH_REQUIRE(nscale >= 1 && nscale <= 4, BLQH_EINVAL, "a shared-scale batch takes 1 to 4 scales");
    for (int g = 0; g < nscale; ++g)
        BLQH_REQUIRE(std::isfinite(scales[g]) && scales[g] >= 0.0, BLQH_EINVAL, "scales must be finite and >= 0");
    auto* b = new blqh_batch;
    try {
        const int G = nscale;
        b->G = G;
        b->tpl = make_templates(base, tun, nsec);
        const int P = nsec * G;
        b->hs.resize(P);
        b->item0.assign(P + 1, 0);
        b->gitem0.assign(nsec + 1, 0);
        b->tscale.resize(P);
        std::vector<int32_t> ip, gp;
        for (int q = 0; q < nsec; ++q) {
            blqh_hamil* H = b->tpl[q];
            const int64_t nit = ceil_div(H->n, kItemRows);
            b->gitem0[q + 1] = b->gitem0[q] + nit;
            for (int64_t i = 0; i < nit; ++i) gp.push_back(q);
            b->nnz += H->nnz;
            for (int g = 0; g < G; ++g) {
                const int p = q * G + g;
                b->hs[p] = H;
                b->tscale[p] = scales[g];
                BProb d{};
                d.n = H->n;
                d.npan = 1;
                d.real = H->pan.real ? 1 : 0;
                d.item0 = b->nitems;
                d.voff = b->rows;
                for (int64_t i = 0; i < nit; ++i) ip.push_back(p);
                b->nitems += nit;
                b->rows += H->n;
                b->item0[p + 1] = b->nitems;
                b->h_probs.push_back(d);
            }
        }
        b->ngitems = b->gitem0[nsec];
        BLQH_REQUIRE(b->nitems < (int64_t)1 << 31, BLQH_ERANGE, "batch too large");
        b->item_prob.alloc(ip.size());
        copy_sync(b->item_prob.get(), ip.data(), ip.size() * 4, hipMemcpyHostToDevice);
        b->gitem_grp.alloc(gp.size());
        copy_sync(b->gitem_grp.get(), gp.data(), gp.size() * 4, hipMemcpyHostToDevice);
        b->dgitem0.alloc(nsec + 1);
        copy_sync(b->dgitem0.get(), b->gitem0.data(), (nsec + 1) * 8, hipMemcpyHostToDevice);
        b->dtscale.alloc(P);
        copy_sync(b->dtscale.get(), b->tscale.data(), P * 8, hipMemcpyHostToDevice);
        b->grid = (int)(8 * ceil_div(std::min<int64_t>(b->nitems, (int64_t)runtime().num_cus * 8), 8));
        b->xitems.alloc(2 * b->nitems);
        b->xseg.alloc(27);
        b->relayout.alloc(panel_layout_scratch_bytes(b->tpl.data(), nsec));
        b->lpr = 64;
        refresh_probs(b);
    } catch (...) {
        delete b;
        throw;
    }
    return b;
}

// The next parameter point of a shared-scale batch: new H0 and T of the same
// sectors (same sizes, same order); the scales stay.
void blqh_batch_rebind_scaled_impl(blqh_batch* b, const blqh_hamil* const* base, const blqh_hamil* const* tun,
                                   int nsec) {
    BLQH_REQUIRE(b->G > 0, BLQH_EINVAL, "not a shared-scale batch");
    BLQH_REQUIRE(nsec == (int)b->tpl.size(), BLQH_EINVAL, "rebind needs as many sectors as the batch has");
    for (int q = 0; q < nsec; ++q)
        BLQH_REQUIRE(base[q] != nullptr && base[q]->n == b->tpl[q]->n, BLQH_EINVAL,
                     "rebind needs sectors of the same sizes, in the same order");
    std::vector<blqh_hamil*> fresh = make_templates(base, tun, nsec);
    BLQH_HIP(hipStreamSynchronize(stream()));  // no queued work reads the old templates
    for (blqh_hamil* H : b->tpl) delete H;
    b->tpl.swap(fresh);
    b->nnz = 0;
    for (int q = 0; q < nsec; ++q) {
        b->nnz += b->tpl[q]->nnz;
        for (int g = 0; g < b->G; ++g) {
            const int p = q * b->G + g;
            b->hs[p] = b->tpl[q];
            b->h_probs[p].real = b->tpl[q]->pan.real ? 1 : 0;
        }
    }
    const size_t scratch = panel_layout_scratch_bytes(b->tpl.data(), nsec);
    if (scratch > b->relayout.bytes()) b->relayout.alloc(scratch);
    refresh_probs(b);
}

int blqh_batch_group_size_impl(const blqh_batch* b) { return b->G; }

void blqh_batch_destroy_impl(blqh_batch* b) { delete b; }

int blqh_batch_lanes_impl(const blqh_batch* b) { return b->lpr; }

void blqh_batch_set_lanes_impl(blqh_batch* b, int lanes) {
    BLQH_REQUIRE(lanes == 4 || lanes == 8 || lanes == 16 || lanes == 32 || lanes == 64, BLQH_EINVAL,
                 "unsupported lanes per row");
    BLQH_REQUIRE(b->G == 0 || lanes == 64, BLQH_EINVAL, "a shared-scale batch runs one lane per row");
    b->lpr = lanes;
}

void blqh_batch_lanczos_impl(blqh_batch* b, const blqh_lanczos_opts& o, const uint32_t* seeds,
                             const double* const* starts, blqh_lanczos_result* res, double* const* psi_out,
                             bool cont, double excited) {
    BLQH_REQUIRE(o.start_mix >= 0 && std::isfinite(o.start_mix), BLQH_EINVAL, "start_mix must be finite and >= 0");
    BLQH_REQUIRE(excited >= 0 && std::isfinite(excited), BLQH_EINVAL, "excited_mix must be finite and >= 0");
    if (cont)
        BLQH_REQUIRE(b->cont_ok && b->pool_cap == o.lambda && !starts, BLQH_EINVAL,
                     "a continuation solve needs a finished solve of the same lambda on this batch");
    b->cont_ok = false;  // the steps below overwrite pl.u
    auto t0 = std::chrono::steady_clock::now();
    // BLQH_BATCH_TIMING=1 prints the host-side phases of every solve to stderr
    const bool phase_log = std::getenv("BLQH_BATCH_TIMING") != nullptr;
    auto tp = t0;
    auto phase = [&](const char* what) {
        if (!phase_log) return;
        BLQH_HIP(hipStreamSynchronize(stream()));
        const auto now = std::chrono::steady_clock::now();
        std::fprintf(stderr, "blqh batch: %-12s %8.2f ms\n", what, std::chrono::duration<double>(now - tp).count() * 1e3);
        tp = now;
    };
    hipStream_t st = stream();
    const int P = (int)b->hs.size();
    const int cap = o.lambda;
    BLQH_REQUIRE(cap >= 2 && cap <= 100000, BLQH_EINVAL, "lambda out of range");
    const int chk = std::max(1, o.check_every);
    const int64_t ld = cap + 1;

    // Panel pointers are re-read here: a single-problem solve or SpMV of a
    // member Hamiltonian may have reordered its panels (new buffers) since
    // the batch was created.
    refresh_probs(b);
    // Krylov blocks of all problems and the per-problem scalars (kept for the
    // next solve with the same lambda)
    Pools& pl = b->pl;
    if (b->pool_cap != cap) {
        if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        b->gexec = nullptr;
        b->graph = nullptr;
        b->pool_cap = -1;
        b->pl = Pools{};
        size_t vecs = 0;
        for (int p = 0; p < P; ++p) vecs += (size_t)ld * b->h_probs[p].n;
        pl.W.alloc(vecs);
        pl.probs.alloc(P);
        pl.active.alloc(P);
        pl.jbase.alloc(1);
        pl.mvec.alloc(P);
        pl.alpha.alloc((size_t)P * cap);
        pl.beta.alloc((size_t)P * ld);
        pl.invb.alloc((size_t)P * ld);
        pl.part1.alloc(b->nitems);
        pl.part2.alloc(b->nitems);
        pl.coef.alloc((size_t)P * cap);
        pl.coefx.alloc((size_t)P * cap);
        pl.xs.alloc(b->rows);
        pl.xref.alloc(b->rows);
        pl.scale.alloc(P);
        pl.u.alloc(b->rows);
        pl.psi.alloc(b->rows);
        if (b->pin) (void)hipHostFree(b->pin);
        b->pin = nullptr;
        // alpha, beta, two slots of active flags, two slots of live work lists
        const size_t pin_bytes = ((size_t)P * cap + (size_t)P * ld) * 8 + 2 * (size_t)P * 4 +
                                 2 * (2 * (size_t)b->nitems + 27) * 4;
        BLQH_HIP(hipHostMalloc((void**)&b->pin, pin_bytes, hipHostMallocDefault));
        b->pool_cap = cap;
    }
    std::vector<BProb> hp = b->h_probs;
    {
        size_t off = 0;
        for (int p = 0; p < P; ++p) {
            hp[p].W = pl.W.get() + off;
            off += (size_t)ld * hp[p].n;
        }
    }
    BLQH_HIP(hipMemcpyAsync(pl.probs.get(), hp.data(), P * sizeof(BProb), hipMemcpyHostToDevice, st));

    phase("setup");
    // start vectors, beta_0 = 1: the caller's on the host (normalised, mixed,
    // permuted), the seeded ones drawn on the device (k_bmt_start)
    if (cont) {
        // conj(previous ground state) + start_mix * seeded unit vector, drawn
        // into the second Krylov slot (free until the first step)
        std::vector<BStart> gj;
        std::vector<BCont> cj(P);
        for (int p = 0; p < P; ++p) {
            const blqh_hamil* Hp = b->hs[p];
            double2* gen = hp[p].W + hp[p].n;
            cj[p] = {hp[p].n, pl.u.get() + hp[p].voff, gen, pl.xref.get() + hp[p].voff, hp[p].W,
                     Hp->ord.on ? Hp->ord.iperm.get() : nullptr, 0.0, excited};
            if (o.start_mix > 0) gj.push_back({hp[p].n, seeds ? seeds[p] : o.seed, gen, gen, nullptr});
        }
        DevBuf<BStart> dgj;
        DevBuf<BCont> dcj(P);
        DevBuf<double> dnrm(2 * (size_t)P);
        std::vector<double> nrm(P);
        if (!gj.empty()) {
            dgj.alloc(P);
            BLQH_HIP(hipMemcpyAsync(dgj.get(), gj.data(), P * sizeof(BStart), hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_bmt_start, dim3(P), dim3(1024), 0, st, dgj.get(), dnrm.get());
            BLQH_HIP(hipGetLastError());
            BLQH_HIP(hipMemcpyAsync(nrm.data(), dnrm.get(), P * sizeof(double), hipMemcpyDeviceToHost, st));
            BLQH_HIP(hipStreamSynchronize(st));
            for (int p = 0; p < P; ++p) cj[p].mix = o.start_mix / std::sqrt(nrm[p]);
        }
        BLQH_HIP(hipMemcpyAsync(dcj.get(), cj.data(), P * sizeof(BCont), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_bcont, dim3(P), dim3(1024), 0, st, dcj.get(), dnrm.get());
        BLQH_HIP(hipGetLastError());
        BLQH_HIP(hipMemcpyAsync(nrm.data(), dnrm.get(), P * sizeof(double), hipMemcpyDeviceToHost, st));
        BLQH_HIP(hipStreamSynchronize(st));
        std::vector<BStart> sj(P);
        for (int p = 0; p < P; ++p) {
            BLQH_REQUIRE(nrm[p] > 0 && std::isfinite(nrm[p]), BLQH_EINVAL, "continuation start is zero or not finite");
            nrm[p] = std::sqrt(nrm[p]);
            sj[p] = {hp[p].n, 0u, hp[p].W, hp[p].W, nullptr};
        }
        BLQH_HIP(hipMemcpyAsync(dnrm.get() + P, nrm.data(), P * sizeof(double), hipMemcpyHostToDevice, st));
        DevBuf<BStart> dsj(P);
        BLQH_HIP(hipMemcpyAsync(dsj.get(), sj.data(), P * sizeof(BStart), hipMemcpyHostToDevice, st));
        int64_t nmax = 0;
        for (int p = 0; p < P; ++p) nmax = std::max(nmax, hp[p].n);
        const unsigned gx = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(nmax, kBT), 256));
        for (int j0 = 0; j0 < P; j0 += 65535)
            hipLaunchKernelGGL(k_bstart_scale, dim3(gx, (unsigned)std::min(65535, P - j0)), dim3(kBT), 0, st,
                               dsj.get() + j0, dnrm.get() + P + j0);
        BLQH_HIP(hipGetLastError());
        std::vector<double> ones((size_t)P * ld, 1.0);
        BLQH_HIP(hipMemcpyAsync(pl.beta.get(), ones.data(), ones.size() * 8, hipMemcpyHostToDevice, st));
        BLQH_HIP(hipMemcpyAsync(pl.invb.get(), ones.data(), ones.size() * 8, hipMemcpyHostToDevice, st));
        BLQH_HIP(hipStreamSynchronize(st));  // job tables, norms and ones are freed at scope exit
    } else {
        std::vector<std::vector<double2>> v0(P);
        std::vector<char> seeded(P, 0);
        for (int p = 0; p < P; ++p) seeded[p] = !(starts && starts[p]);
        auto make = [&](int p0, int p1) {
            for (int p = p0; p < p1; ++p) {
                if (seeded[p]) continue;
                const int64_t n = hp[p].n;
                v0[p].assign((const double2*)starts[p], (const double2*)starts[p] + n);
                double s = 0;
                for (auto& z : v0[p]) s += z.x * z.x + z.y * z.y;
                if (!(s > 0) || !std::isfinite(s)) { v0[p].clear(); continue; }
                if (o.start_mix > 0) {  // start/|start| + start_mix * seeded unit vector
                    std::vector<double2> r;
                    start_vector(n, seeds ? seeds[p] : o.seed, r);
                    const double zs = 1.0 / std::sqrt(s);
                    s = 0;
                    for (int64_t i = 0; i < n; ++i) {
                        double2& z = v0[p][i];
                        z = make_double2(z.x * zs + o.start_mix * r[i].x, z.y * zs + o.start_mix * r[i].y);
                        s += z.x * z.x + z.y * z.y;
                    }
                }
                s = std::sqrt(s);
                for (auto& z : v0[p]) { z.x /= s; z.y /= s; }
                const blqh_hamil* Hp = b->hs[p];
                if (Hp->ord.on) {  // the member's SpMV formats are in its internal order
                    std::vector<double2> w(n);
                    for (int64_t r = 0; r < n; ++r) w[r] = v0[p][Hp->ord.h_iperm[r]];
                    v0[p].swap(w);
                }
            }
        };
        int ncaller = 0;
        for (int p = 0; p < P; ++p) ncaller += seeded[p] ? 0 : 1;
        if (ncaller > 0) {
            const int nt = std::max(1, std::min<int>(ncaller, (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()))));
            std::vector<std::thread> th;
            for (int t = 0; t < nt; ++t)
                th.emplace_back(make, (int)((int64_t)P * t / nt), (int)((int64_t)P * (t + 1) / nt));
            for (auto& t : th) t.join();
        }
        for (int p = 0; p < P; ++p)
            BLQH_REQUIRE(seeded[p] || !v0[p].empty(), BLQH_EINVAL, "start vector is zero or not finite");
        for (int p = 0; p < P; ++p)
            if (!seeded[p])
                BLQH_HIP(hipMemcpyAsync(hp[p].W, v0[p].data(), hp[p].n * sizeof(double2), hipMemcpyHostToDevice, st));
        std::vector<BStart> jobs;
        int64_t nmax = 0;
        for (int p = 0; p < P; ++p)
            if (seeded[p]) {
                const blqh_hamil* Hp = b->hs[p];
                const bool perm = Hp->ord.on;
                jobs.push_back({hp[p].n, seeds ? seeds[p] : o.seed, perm ? pl.u.get() + hp[p].voff : hp[p].W, hp[p].W,
                                perm ? Hp->ord.iperm.get() : nullptr});
                nmax = std::max(nmax, hp[p].n);
            }
        DevBuf<BStart> djobs;
        DevBuf<double> dnrm;
        if (!jobs.empty()) {
            const int J = (int)jobs.size();
            djobs.alloc(J);
            dnrm.alloc(2 * (size_t)J);
            BLQH_HIP(hipMemcpyAsync(djobs.get(), jobs.data(), J * sizeof(BStart), hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_bmt_start, dim3(J), dim3(1024), 0, st, djobs.get(), dnrm.get());
            BLQH_HIP(hipGetLastError());
            std::vector<double> nrm(J);
            BLQH_HIP(hipMemcpyAsync(nrm.data(), dnrm.get(), J * sizeof(double), hipMemcpyDeviceToHost, st));
            BLQH_HIP(hipStreamSynchronize(st));
            for (double& x : nrm) x = std::sqrt(x);
            BLQH_HIP(hipMemcpyAsync(dnrm.get() + J, nrm.data(), J * sizeof(double), hipMemcpyHostToDevice, st));
            const unsigned gx = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(nmax, kBT), 256));
            for (int j0 = 0; j0 < J; j0 += 65535)
                hipLaunchKernelGGL(k_bstart_scale, dim3(gx, (unsigned)std::min(65535, J - j0)), dim3(kBT), 0, st,
                                   djobs.get() + j0, dnrm.get() + J + j0);
            BLQH_HIP(hipGetLastError());
        }
        std::vector<double> ones((size_t)P * ld, 1.0);
        BLQH_HIP(hipMemcpyAsync(pl.beta.get(), ones.data(), ones.size() * 8, hipMemcpyHostToDevice, st));
        BLQH_HIP(hipMemcpyAsync(pl.invb.get(), ones.data(), ones.size() * 8, hipMemcpyHostToDevice, st));
        BLQH_HIP(hipStreamSynchronize(st));  // v0, ones, the job table and the norms are freed at scope exit
    }
    phase("start");
    std::vector<int32_t> act(P, 1);
    BLQH_HIP(hipMemcpyAsync(pl.active.get(), act.data(), P * 4, hipMemcpyHostToDevice, st));
    BState S = make_state(b, pl, cap);

    // `chk` steps as one graph; kernels read the first step from pl.jbase and
    // the graph's last kernel advances it
    Profile& prof = profile();
    const bool timed = prof.on;  // per-kernel events need plain launches
    std::vector<hipEvent_t> ev;
    int ntimed = 0;
    auto chunk = [&](bool record, int first) {
        for (int k = 0; k < chk; ++k) {
            const bool t = record && first + k < cap;
            if (t) {
                while ((int)ev.size() < 2 * (ntimed + 1)) {
                    hipEvent_t e;
                    BLQH_HIP(hipEventCreate(&e));
                    ev.push_back(e);
                }
                BLQH_HIP(hipEventRecord(ev[2 * ntimed], st));
            }
            launch_b1<true>(b, S, k, nullptr, nullptr, pl.part1.get(), st);
            if (t) BLQH_HIP(hipEventRecord(ev[2 * ntimed++ + 1], st));
            hipLaunchKernelGGL(k_b2, dim3(b->grid), dim3(kBT), 0, st, S, k);
            hipLaunchKernelGGL(k_b3, dim3(P), dim3(64), 0, st, S, k, P);
        }
        hipLaunchKernelGGL(k_bnext, dim3(1), dim3(64), 0, st, pl.jbase.get(), chk);
    };
    struct EvList {
        std::vector<hipEvent_t>& ev;
        ~EvList() {
            for (auto x : ev) (void)hipEventDestroy(x);
        }
    } evguard{ev};
    const int kinds = (b->has_real ? 2 : 0) | (b->has_complex ? 1 : 0);  // which K1 launches a step has
    if (!timed && (b->g_chk != chk || b->g_lpr != b->lpr || b->g_cap != cap || b->g_kinds != kinds || !b->gexec)) {
        if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        b->gexec = nullptr;
        b->graph = nullptr;
        BLQH_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        chunk(false, 0);
        BLQH_HIP(hipStreamEndCapture(st, &b->graph));
        BLQH_HIP(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
        b->g_chk = chk;
        b->g_lpr = b->lpr;
        b->g_cap = cap;
        b->g_kinds = kinds;
    }
    hipGraphExec_t gexec = b->gexec;
    const int zero = 0;
    copy_sync(pl.jbase.get(), &zero, 4, hipMemcpyHostToDevice);

    // pinned staging: alpha/beta read back per replay, active flags (two slots,
    // so a slot is rewritten only after its copy has run)
    char* pin = b->pin;
    double* ha = (double*)pin;
    double* hb = ha + (size_t)P * cap;
    int32_t* actpin = (int32_t*)(hb + (size_t)P * ld);
    // Work lists of the live members only, rebuilt whenever problems
    // converge: the lists deal whole members to XCDs, so without a rebuild
    // the last live members would sit on a few XCDs while the others idle.
    // Every item's sums keep their owner and order, so results do not depend
    // on the lists.  Two pinned slots, as for the active flags.
    int32_t* listpin = actpin + 2 * (size_t)P;
    const size_t list_slot = 2 * (size_t)b->nitems + 27;
    int lslot = 0;
    bool lists_live = false;
    hipEvent_t evc;
    BLQH_HIP(hipEventCreateWithFlags(&evc, hipEventDisableTiming));
    struct EvGuard {
        hipEvent_t e;
        ~EvGuard() { (void)hipEventDestroy(e); }
    } eguard{evc};

    std::vector<std::vector<double>> ritz(P);
    std::vector<int> m(P, 0), steps_run(P, 0);
    std::vector<char> conv(P, 0), live(P, 1);
    std::vector<double> dE(P, INFINITY), resid(P, INFINITY);
    // host threads for the convergence tests: worth it only for many problems
    const int nthr = P >= 32 ? (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency())) : 1;
    if (nthr > 1 && (!b->hpool || b->hpool->size() != nthr)) b->hpool.reset(new HostPool(nthr - 1));
    std::vector<std::vector<double>> svecs(nthr, std::vector<double>(cap));
    std::vector<int> live_ids;
    live_ids.reserve(P);
    // The next replays are queued before the host tests the last ones, so the
    // host test overlaps the GPU; a problem that converged in the tested
    // replays runs the queued ones too, whose steps are never used.  While
    // many problems are live one replay is queued ahead; with a quarter or
    // fewer a replay is short (a few tens of microseconds per step) and the
    // host test of each replay would leave the GPU idle, so two, then four
    // replays go ahead of each test.
    int queued = 0;  // steps queued so far
    auto enqueue = [&](int first) {
        if (timed) chunk(true, first);
        else BLQH_HIP(hipGraphLaunch(gexec, st));
        BLQH_HIP(hipGetLastError());
    };
    auto enqueue_n = [&](int r) {
        for (int i = 0; i < r && queued < cap; ++i) {
            enqueue(queued);
            queued += chk;
        }
    };
    auto readback = [&]() {
        BLQH_HIP(hipMemcpyAsync(ha, pl.alpha.get(), (size_t)P * cap * 8, hipMemcpyDeviceToHost, st));
        BLQH_HIP(hipMemcpyAsync(hb, pl.beta.get(), (size_t)P * ld * 8, hipMemcpyDeviceToHost, st));
        BLQH_HIP(hipEventRecord(evc, st));
    };
    int done = 0, nlive = P, slot = 0;
    enqueue_n(1);
    readback();
    while (true) {
        const int upto = std::min(cap, queued);  // steps the last readback covers
        const bool more = queued < cap;
        if (more) enqueue_n(4 * nlive > P ? 1 : 16 * nlive > P ? 2 : 4);  // speculative next replays
        BLQH_HIP(hipEventSynchronize(evc));
        // the single-problem solver's convergence test, per live problem
        // (problems split over host threads; each writes only its own entries)
        auto test = [&](int p, std::vector<double>& svec) {
            steps_run[p] = std::min(cap, queued);
            const double* a = ha + (size_t)p * cap;
            const double* bb = hb + (size_t)p * ld;
            for (int j = done; j < upto; ++j) {
                const std::vector<double>& rz = ritz[p];
                const double e = j == 0 ? tridiag_lowest(1, a, bb + 1, nullptr)
                                        : tridiag_lowest_next(j + 1, a, bb + 1, rz[j - 1],
                                                              j > 1 ? std::fabs(rz[j - 1] - rz[j - 2]) : INFINITY);
                ritz[p].push_back(e);
                m[p] = j + 1;
                if (j > 0) dE[p] = std::fabs(ritz[p][j] - ritz[p][j - 1]);
                const double scale = std::max(std::fabs(a[j]), 1.0);
                if (!(bb[j + 1] > 1e-13 * scale) || !std::isfinite(bb[j + 1])) {
                    resid[p] = 0;
                    conv[p] = 1;
                    break;
                }
                if (o.tol > 0 && j > 0 && dE[p] < o.tol) {
                    if (o.res_tol <= 0) { conv[p] = 1; break; }
                    (void)tridiag_lowest(j + 1, a, bb + 1, svec.data());
                    resid[p] = bb[j + 1] * std::fabs(svec[j]);
                    if (resid[p] < o.res_tol) { conv[p] = 1; break; }
                }
            }
        };
        {
            live_ids.clear();
            for (int p = 0; p < P; ++p)
                if (live[p]) live_ids.push_back(p);
            const int nl = (int)live_ids.size();
            const int nt = std::min(nthr, (nl + 3) / 4);  // at least four problems per thread
            if (nt <= 1) {
                for (int p : live_ids) test(p, svecs[0]);
            } else {
                b->hpool->run(nt, [&](int t) {
                    for (int i = t; i < nl; i += nt) test(live_ids[i], svecs[t]);
                });
            }
        }
        bool changed = false;
        for (int p = 0; p < P; ++p)
            if (live[p] && conv[p]) {
                live[p] = 0;
                --nlive;
                changed = true;
            }
        done = upto;
        if (!more || nlive == 0) break;
        if (changed) {
            int32_t* dst = actpin + (size_t)slot * P;
            for (int p = 0; p < P; ++p) dst[p] = live[p];
            BLQH_HIP(hipMemcpyAsync(pl.active.get(), dst, P * 4, hipMemcpyHostToDevice, st));
            slot ^= 1;
            if (b->lpr == 64) {  // k_b1j walks the lists
                std::vector<int32_t> items, seg;
                xcd_lists(xcd_members(b, &live), items, seg);
                int32_t* lp = listpin + (size_t)lslot * list_slot;
                std::copy(items.begin(), items.end(), lp);
                std::copy(seg.begin(), seg.end(), lp + 2 * b->nitems);
                BLQH_HIP(hipMemcpyAsync(b->xitems.get(), lp, items.size() * 4, hipMemcpyHostToDevice, st));
                BLQH_HIP(hipMemcpyAsync(b->xseg.get(), lp + 2 * b->nitems, 27 * 4, hipMemcpyHostToDevice, st));
                lslot ^= 1;
                lists_live = true;
            }
        }
        readback();
    }
    BLQH_HIP(hipStreamSynchronize(st));  // a speculative replay may still run
    if (lists_live) {  // every member takes part in the final products again
        copy_sync(b->xitems.get(), b->h_xitems.data(), b->h_xitems.size() * 4, hipMemcpyHostToDevice);
        copy_sync(b->xseg.get(), b->h_xseg.data(), 27 * 4, hipMemcpyHostToDevice);
    }

    phase("iterate");
    // pass 2 per problem, then psi = sum_l c_l conj(v_l), normalised, and Re<psi|H psi>
    std::vector<double> hinvb((size_t)P * ld);
    BLQH_HIP(hipMemcpyAsync(hinvb.data(), pl.invb.get(), hinvb.size() * 8, hipMemcpyDeviceToHost, st));
    BLQH_HIP(hipStreamSynchronize(st));
    std::vector<double> w((size_t)P * cap, 0.0), wx((size_t)P * cap, 0.0), E0(P);
    std::vector<int32_t> hm(P);
    for (int p = 0; p < P; ++p)
        BLQH_REQUIRE(m[p] >= 1 && std::isfinite(ritz[p].back()), BLQH_EINTERNAL,
                     "Lanczos produced non-finite coefficients");
    auto pass2 = [&](int p, std::vector<double>& c) {
        c.resize(m[p]);
        E0[p] = tridiag_lowest(m[p], ha + (size_t)p * cap, hb + (size_t)p * ld + 1, c.data());
        for (int l = 0; l < m[p]; ++l) w[(size_t)p * cap + l] = c[l] * hinvb[(size_t)p * ld + l];
        excited_mix(m[p], ha + (size_t)p * cap, hb + (size_t)p * ld + 1, c.data(), wx.data() + (size_t)p * cap);
        for (int l = 0; l < m[p]; ++l) wx[(size_t)p * cap + l] *= hinvb[(size_t)p * ld + l];
        hm[p] = m[p];
        if (!std::isfinite(resid[p])) resid[p] = hb[(size_t)p * ld + m[p]] * std::fabs(c[m[p] - 1]);
    };
    {
        const int nt = std::min(nthr, (P + 3) / 4);  // per-problem results only: any split gives the same values
        if (nt <= 1) {
            for (int p = 0; p < P; ++p) pass2(p, svecs[0]);
        } else {
            b->hpool->run(nt, [&](int t) {
                for (int p = t; p < P; p += nt) pass2(p, svecs[t]);
            });
        }
    }
    BLQH_HIP(hipMemcpyAsync(pl.coef.get(), w.data(), w.size() * 8, hipMemcpyHostToDevice, st));
    BLQH_HIP(hipMemcpyAsync(pl.mvec.get(), hm.data(), P * 4, hipMemcpyHostToDevice, st));
    std::fill(act.begin(), act.end(), 1);
    BLQH_HIP(hipMemcpyAsync(pl.active.get(), act.data(), P * 4, hipMemcpyHostToDevice, st));
    BLQH_HIP(hipMemcpyAsync(pl.coefx.get(), wx.data(), wx.size() * 8, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_bcombine, dim3(b->grid), dim3(kBT), 0, st, S, pl.coef.get(), pl.coefx.get(), pl.mvec.get(),
                       pl.psi.get(), pl.xs.get());
    std::vector<double> hp1(b->nitems), hp2(b->nitems), sc(P);
    BLQH_HIP(hipMemcpyAsync(hp1.data(), pl.part1.get(), b->nitems * 8, hipMemcpyDeviceToHost, st));
    BLQH_HIP(hipStreamSynchronize(st));
    for (int p = 0; p < P; ++p) {
        double s = 0;
        for (int64_t i = b->item0[p]; i < b->item0[p + 1]; ++i) s += hp1[i];
        sc[p] = s > 0 ? 1.0 / std::sqrt(s) : 0.0;
    }
    BLQH_HIP(hipMemcpyAsync(pl.scale.get(), sc.data(), P * 8, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_bscale, dim3(b->grid), dim3(kBT), 0, st, S, pl.scale.get(), pl.psi.get());
    launch_b1<false>(b, S, 0, pl.psi.get(), pl.u.get(), pl.part2.get(), st);
    BLQH_HIP(hipGetLastError());
    BLQH_HIP(hipMemcpyAsync(hp2.data(), pl.part2.get(), b->nitems * 8, hipMemcpyDeviceToHost, st));
    // ground states: all of them in the reference order into u (free after
    // the last SpMV; the start of a continuation solve), then, when any is
    // wanted, one copy of u into pinned staging and host copies into the
    // caller's arrays on the pool threads
    std::vector<BPsi> pj;
    int64_t pmax = 0;
    bool want = false;
    for (int p = 0; p < P; ++p) {
        const blqh_hamil* Hp = b->hs[p];
        pj.push_back({hp[p].n, Hp->ord.on ? Hp->ord.iperm.get() : nullptr, pl.psi.get() + hp[p].voff,
                      pl.u.get() + hp[p].voff});
        pj.push_back({hp[p].n, Hp->ord.on ? Hp->ord.iperm.get() : nullptr, pl.xs.get() + hp[p].voff,
                      pl.xref.get() + hp[p].voff});
        pmax = std::max(pmax, hp[p].n);
        want = want || (psi_out && psi_out[p]);
    }
    DevBuf<BPsi> dpj(pj.size());
    BLQH_HIP(hipMemcpyAsync(dpj.get(), pj.data(), pj.size() * sizeof(BPsi), hipMemcpyHostToDevice, st));
    {
        const unsigned gx = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(pmax, kBT), 256));
        for (size_t j0 = 0; j0 < pj.size(); j0 += 65535)
            hipLaunchKernelGGL(k_bpsi_ref, dim3(gx, (unsigned)std::min<size_t>(65535, pj.size() - j0)), dim3(kBT), 0,
                               st, dpj.get() + j0);
        BLQH_HIP(hipGetLastError());
    }
    if (want) {
        if (!b->psi_pin) BLQH_HIP(hipHostMalloc((void**)&b->psi_pin, (size_t)b->rows * sizeof(double2), hipHostMallocDefault));
        BLQH_HIP(hipMemcpyAsync(b->psi_pin, pl.u.get(), (size_t)b->rows * sizeof(double2), hipMemcpyDeviceToHost, st));
    }
    BLQH_HIP(hipStreamSynchronize(st));
    b->cont_ok = true;
    if (want) {
        auto put = [&](int p) {
            if (psi_out[p]) std::memcpy(psi_out[p], b->psi_pin + hp[p].voff, hp[p].n * sizeof(double2));
        };
        const int nt = std::min(nthr, (P + 3) / 4);
        if (nt <= 1) {
            for (int p = 0; p < P; ++p) put(p);
        } else {
            b->hpool->run(nt, [&](int t) {
                for (int p = t; p < P; p += nt) put(p);
            });
        }
    }
    if (timed) {
        for (int k = 0; k < ntimed; ++k) {
            float ms = 0;
            BLQH_HIP(hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]));
            prof.ms += ms;
        }
        prof.launches += ntimed;
    }
    phase("finish");
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (int p = 0; p < P; ++p) {
        double erq = 0;
        for (int64_t i = b->item0[p]; i < b->item0[p + 1]; ++i) erq += hp2[i];
        // bytes of every K1 launch this problem took part in (speculative steps included)
        // (a shared-scale group reads its matrix once for its G problems)
        if (timed)
            prof.bytes += (double)steps_run[p] * (20.0 * b->hs[p]->nnz / (b->G ? b->G : 1) + 56.0 * hp[p].n + 8.0);
        blqh_lanczos_result& r = res[p];
        std::memset(&r, 0, sizeof(r));
        r.E0 = E0[p];
        r.E_rq = erq;
        r.dE = dE[p];
        r.iterations = m[p];
        r.converged = conv[p];
        r.passes = 2;
        r.seconds = secs;
        r.residual = resid[p];
    }
}
