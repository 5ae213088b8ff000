// lanczos.hip -- complex FP64 CSR SpMV and the Lanczos ground-state solver.
//
// Reference: Mat::operator* (BLQH_without_spin/matrix.cpp:244-267),
// coeff_explicit_update / diag / eigenstates_reconstruction /
// ground_state_energy (BLQH_without_spin/lanczos_hamiltonian.cpp:453-656).
//
// Vectors are kept unnormalised: v_j = W_j / beta_j.  One Lanczos step is two
// kernels:
//   K1  u = (H W_j)/beta_j - beta_j W_{j-1}/beta_{j-1},  alpha_j = Re<v_j|u>
//   K2  W_{j+1} = u - alpha_j v_j,  beta_{j+1} = |W_{j+1}|
// Both reduce through per-block partials summed in block order by the last
// block to finish, so every scalar is bit-reproducible: a replayed pass 3
// regenerates exactly the pass-1 vectors.  When the Krylov vectors fit in
// HBM they are kept (W_j are written in place) and the ground state is one
// weighted sum over them instead of a replay.
// The reconstructed state is psi = sum_l c_l conj(v_l) as in the reference
// (lanczos_hamiltonian.cpp:624-640), normalised; E_rq = Re<psi|H psi>.
#include <algorithm>
#include <chrono>
#include <mutex>
#include <cmath>
#include <cstring>
#include <cstdio>
#include <cstdlib>
#include <random>

#include "engine.hpp"
#include "mtstart.hpp"

namespace {

using namespace blqh;

constexpr int kThreads = 256;

typedef double d2v __attribute__((ext_vector_type(2)));
// streaming (non-temporal) load of a double2: matrix data is read once per
// SpMV and should not evict the gathered vector from L2 / MALL
__device__ __forceinline__ double2 ntload(const double2* p) {
    const d2v v = __builtin_nontemporal_load(reinterpret_cast<const d2v*>(p));
    return make_double2(v.x, v.y);
}

// ---- deterministic grid reduction -------------------------------------------
// Block sum of v -> partials[blockIdx].  The grid total is formed by the
// next kernel on the stream (wave_sum, fixed order): no atomics or
// device-scope fences, whose cost grows with the number of workgroups (a
// ticket taken by each of 4,700 workgroups made the fused step of an
// 18.8k-row matrix ten times slower than its product).
__device__ __forceinline__ void block_partial(double v, double* partials) {
    __shared__ double sh[16];  // up to 1024 threads
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += sh[w];
        partials[blockIdx.x] = t;
    }
}

// sum of part[0 .. cnt) by the whole workgroup in a fixed order, returned
// to every thread: each thread adds its strided share with four independent
// loads in flight (thousands of partials are one memory latency, not a chain
// of them), then the wavefronts' butterflies and a fixed-order sum of the
// wavefront totals.  Contains __syncthreads: call from all threads.
__device__ __forceinline__ double block_total(const double* __restrict__ part, int cnt) {
    __shared__ double sh[17];
    const int nt = (int)blockDim.x;
    double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    int i = threadIdx.x;
    for (; i + 3 * nt < cnt; i += 4 * nt) {
        a0 += part[i];
        a1 += part[i + nt];
        a2 += part[i + 2 * nt];
        a3 += part[i + 3 * nt];
    }
    for (; i < cnt; i += nt) a0 += part[i];
    double v = (a0 + a1) + (a2 + a3);
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    __syncthreads();  // sh may still be read by a previous call
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int w = 0; w < nt / 64; ++w) t += sh[w];
        sh[16] = t;
    }
    __syncthreads();
    return sh[16];
}

struct Scal {  // device scalars per iteration
    double* alpha;  // [cap]
    double* beta;   // [cap+1], beta[0] = 1
    double* invb;   // [cap+1], invb[0] = 1
};

// ---- SpMV ---------------------------------------------------------------------
template <int LPR>
__device__ __forceinline__ double2 row_dot(const int64_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                           const double2* __restrict__ val, const double2* __restrict__ x,
                                           int64_t row, int lane) {
    const int64_t s = rowptr[row], e = rowptr[row + 1];
    double re = 0.0, im = 0.0;
    for (int64_t k = s + lane; k < e; k += LPR) {
        const double2 v = ntload(&val[k]);
        const int c = __builtin_nontemporal_load(&col[k]);
        const double2 xv = x[c];
        re += v.x * xv.x - v.y * xv.y;
        im += v.x * xv.y + v.y * xv.x;
    }
#pragma unroll
    for (int o = LPR / 2; o > 0; o >>= 1) {
        re += __shfl_xor(re, o, LPR);
        im += __shfl_xor(im, o, LPR);
    }
    return make_double2(re, im);
}

template <int LPR>
__global__ void __launch_bounds__(kThreads) k_spmv(int64_t n, const int64_t* __restrict__ rowptr,
                                                    const int32_t* __restrict__ col, const double2* __restrict__ val,
                                                    const double2* __restrict__ x, double2* __restrict__ y) {
    constexpr int RPB = kThreads / LPR;
    const int lane = threadIdx.x % LPR;
    for (int64_t row = (int64_t)blockIdx.x * RPB + threadIdx.x / LPR; row < n; row += (int64_t)gridDim.x * RPB) {
        const double2 r = row_dot<LPR>(rowptr, col, val, x, row, lane);
        if (lane == 0) y[row] = r;
    }
}

// K1: u = invb[j] (H W) - (j > 0 ? beta[j] invb[j-1] Wp : 0); alpha[j] = invb[j] Re<W|u>
template <int LPR>
__global__ void __launch_bounds__(kThreads) k_lz1(int64_t n, const int64_t* __restrict__ rowptr,
                                                   const int32_t* __restrict__ col, const double2* __restrict__ val,
                                                   const double2* __restrict__ W, const double2* __restrict__ Wp,
                                                   double2* __restrict__ u, Scal sc, int j, double* partials) {
    constexpr int RPB = kThreads / LPR;
    const int lane = threadIdx.x % LPR;
    const double ib = sc.invb[j];
    const double cp = j > 0 ? sc.beta[j] * sc.invb[j - 1] : 0.0;
    double acc = 0.0;
    for (int64_t row = (int64_t)blockIdx.x * RPB + threadIdx.x / LPR; row < n; row += (int64_t)gridDim.x * RPB) {
        const double2 r = row_dot<LPR>(rowptr, col, val, W, row, lane);
        if (lane == 0) {
            double ur = ib * r.x, ui = ib * r.y;
            if (j > 0) {
                const double2 p = Wp[row];
                ur -= cp * p.x;
                ui -= cp * p.y;
            }
            u[row] = make_double2(ur, ui);
            const double2 w = W[row];
            acc += w.x * ur + w.y * ui;
        }
    }
    block_partial(acc, partials);
}

// K2: alpha[j] = invb[j] (sum of K1's npart1 partials); Wn = u - alpha[j]
// invb[j] W; partials of |Wn|^2; for s < ns: psi_s += coef[s*ldc + j] invb[j]
// conj(W)   (psi_s = psi + s*n)
__global__ void __launch_bounds__(kThreads) k_lz2(int64_t n, const double2* __restrict__ u,
                                                   const double2* __restrict__ W, double2* __restrict__ Wn, Scal sc,
                                                   int j, int ns, double2* __restrict__ psi,
                                                   const double* __restrict__ coef, int ldc,
                                                   const double* __restrict__ part1, int npart1, double* part2) {
    const double alpha = block_total(part1, npart1) * sc.invb[j];
    if (blockIdx.x == 0 && threadIdx.x == 0) sc.alpha[j] = alpha;
    const double a = alpha * sc.invb[j];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        const double2 w = W[i], x = u[i];
        const double2 r = make_double2(x.x - a * w.x, x.y - a * w.y);
        Wn[i] = r;
        acc += r.x * r.x + r.y * r.y;
        for (int s = 0; s < ns; ++s) {
            const double cj = coef[(int64_t)s * ldc + j] * sc.invb[j];
            double2 p = psi[(int64_t)s * n + i];
            p.x += cj * w.x;
            p.y -= cj * w.y;
            psi[(int64_t)s * n + i] = p;
        }
    }
    block_partial(acc, part2);
}

// K3: beta[j+1] = sqrt(sum of K2's partials), invb[j+1]; one workgroup
__global__ void __launch_bounds__(kThreads) k_lz3(Scal sc, int j, const double* __restrict__ part2, int npart2) {
    const double t = block_total(part2, npart2);
    if (threadIdx.x == 0) {
        const double b = sqrt(t);
        sc.beta[j + 1] = b;
        sc.invb[j + 1] = 1.0 / b;
    }
}

// *out = sum of part[0 .. cnt); one workgroup
__global__ void __launch_bounds__(kThreads) k_psum(const double* __restrict__ part, int cnt, double* out) {
    const double t = block_total(part, cnt);
    if (threadIdx.x == 0) *out = t;
}

// psi += coef[l] invb[l] conj(W_l) for one vector (last replayed vector)
__global__ void k_accum(int64_t n, const double2* __restrict__ W, double c, double2* __restrict__ psi) {
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        const double2 w = W[i];
        double2 p = psi[i];
        p.x += c * w.x;
        p.y -= c * w.y;
        psi[i] = p;
    }
}

// psi = sum_l coef[l] conj(W_l), W_l stored with stride ld
__global__ void __launch_bounds__(kThreads) k_combine(int64_t n, int m, const double2* __restrict__ Wb, int64_t ld,
                                                       const double* __restrict__ coef, double2* __restrict__ psi) {
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        double re = 0.0, im = 0.0;
        for (int l = 0; l < m; ++l) {
            const double2 w = ntload(&Wb[(int64_t)l * ld + i]);
            re += coef[l] * w.x;
            im -= coef[l] * w.y;
        }
        psi[i] = make_double2(re, im);
    }
}

__global__ void __launch_bounds__(kThreads) k_norm2(int64_t n, const double2* __restrict__ x, double* partials) {
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        const double2 v = x[i];
        acc += v.x * v.x + v.y * v.y;
    }
    block_partial(acc, partials);
}

__global__ void k_scale(int64_t n, double2* __restrict__ x, const double* __restrict__ nrm2) {
    const double s = 1.0 / sqrt(*nrm2);
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        double2 v = x[i];
        v.x *= s;
        v.y *= s;
        x[i] = v;
    }
}

// Re<x|y>
__global__ void __launch_bounds__(kThreads) k_redot(int64_t n, const double2* __restrict__ x,
                                                     const double2* __restrict__ y, double* partials) {
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        const double2 a = x[i], b = y[i];
        acc += a.x * b.x + a.y * b.y;
    }
    block_partial(acc, partials);
}

__global__ void __launch_bounds__(kThreads) k_copy(const double2* __restrict__ src, double2* __restrict__ dst,
                                                    int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
        dst[i] = src[i];
}

// ---- panel-major SpMV ---------------------------------------------------------------
// Persistent kernel: workgroup b owns rows [b*rows_per_wg, ...), walks the panels
// in order with its rows' partial sums in LDS (chunk rows at a time), then runs
// the epilogue.  LZ: the fused Lanczos K1 epilogue, x = W.
__device__ __forceinline__ double ntload1(const double* p) { return __builtin_nontemporal_load(p); }

template <bool REAL>
__device__ __forceinline__ void seg_fma(const PanelView& P, const double2* __restrict__ xp, int64_t k, double& re,
                                        double& im) {
    const int c = __builtin_nontemporal_load(&P.pcol[k]);
    if (REAL) {
        const double v = ntload1(&P.pvr[k]);
        const double2 xv = xp[c];
        re += v * xv.x;
        im += v * xv.y;
    } else {
        const double2 v = ntload(&P.pvc[k]);
        const double2 xv = xp[c];
        re += v.x * xv.x - v.y * xv.y;
        im += v.x * xv.y + v.y * xv.x;
    }
}

// U rows per lane group at a time (independent loads in flight), the next
// group's row offsets prefetched while the current one computes.
template <int LPR, int PT, int U, bool REAL, bool LZ>
__global__ void __launch_bounds__(PT) k_pspmv(PanelView P, const double2* __restrict__ x, double2* __restrict__ out,
                                              int64_t rows_per_wg, int chunk, const double2* __restrict__ Wp, Scal sc,
                                              int j, double* partials) {
    extern __shared__ double2 yl[];
    constexpr int NSUB = PT / LPR;
    const int tid = threadIdx.x, sub = tid / LPR, lane = tid % LPR;
    const int64_t n = P.n;
    const int64_t r0 = (int64_t)blockIdx.x * rows_per_wg;
    const int64_t r1 = r0 + rows_per_wg < n ? r0 + rows_per_wg : n;
    double ib = 0, cp = 0;
    if (LZ) {
        ib = sc.invb[j];
        cp = j > 0 ? sc.beta[j] * sc.invb[j - 1] : 0.0;
    }
    double dacc = 0.0;
    for (int64_t c0 = r0; c0 < r1; c0 += chunk) {
        const int nr = (int)(r1 - c0 < chunk ? r1 - c0 : chunk);
        for (int i = tid; i < nr; i += PT) yl[i] = make_double2(0.0, 0.0);
        __syncthreads();
        for (int p = 0; p < P.npan; ++p) {
            const uint32_t* rp = P.prow + (int64_t)p * (n + 1) + c0;
            const int64_t base = P.pbase[p];
            const double2* xp = x + (int64_t)p * kPanelWidth;
            int rb = sub * U;
            uint32_t s[U], e[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int r = rb + u;
                s[u] = r < nr ? rp[r] : 0u;
                e[u] = r < nr ? rp[r + 1] : 0u;
            }
            while (rb < nr) {
                const int nb = rb + NSUB * U;
                uint32_t sn[U], en[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int r = nb + u;
                    sn[u] = r < nr ? rp[r] : 0u;
                    en[u] = r < nr ? rp[r + 1] : 0u;
                }
                double re[U], im[U];
                uint32_t mx = 0;
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    re[u] = 0.0;
                    im[u] = 0.0;
                    mx = e[u] - s[u] > mx ? e[u] - s[u] : mx;
                }
                for (uint32_t off = lane; off < mx; off += LPR) {
#pragma unroll
                    for (int u = 0; u < U; ++u)
                        if (s[u] + off < e[u]) seg_fma<REAL>(P, xp, base + s[u] + off, re[u], im[u]);
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
#pragma unroll
                    for (int o = LPR / 2; o > 0; o >>= 1) {
                        re[u] += __shfl_xor(re[u], o, LPR);
                        im[u] += __shfl_xor(im[u], o, LPR);
                    }
                }
                if (lane == 0) {
#pragma unroll
                    for (int u = 0; u < U; ++u)
                        if (rb + u < nr) {
                            double2 t = yl[rb + u];
                            t.x += re[u];
                            t.y += im[u];
                            yl[rb + u] = t;
                        }
                }
                rb = nb;
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    s[u] = sn[u];
                    e[u] = en[u];
                }
            }
        }
        __syncthreads();
        for (int i = tid; i < nr; i += PT) {
            const int64_t row = c0 + i;
            const double2 yv = yl[i];
            if (LZ) {
                double ur = ib * yv.x, ui = ib * yv.y;
                if (j > 0) {
                    const double2 q = Wp[row];
                    ur -= cp * q.x;
                    ui -= cp * q.y;
                }
                out[row] = make_double2(ur, ui);
                const double2 w = x[row];
                dacc += w.x * ur + w.y * ui;
            } else {
                out[row] = yv;
            }
        }
        __syncthreads();
    }
    if (LZ) block_partial(dacc, partials);
}

// ---- slice-streaming panel SpMV ------------------------------------------------------
// A wavefront owns a slice of 64 consecutive rows (lane l <-> row l) and keeps
// their sums in registers across all panels.  Within a panel the slice's
// entries are one contiguous run; it is streamed in windows of 64 entries (one
// per lane, fully coalesced), the products are prefix-summed across the
// wavefront, and each row lane adds P[last] - P[first-1] of its own segment's
// part of the window.  Workgroups of WPB wavefronts sweep the panels together
// (one barrier per panel) and the persistent grid walks slice groups in the
// same order, so the active x panel stays resident in L2.
__device__ __forceinline__ double wave_incl_scan(double v, int lane) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const double u = __shfl_up(v, d, 64);
        if (lane >= d) v += u;
    }
    return v;
}

// Wavefront `wid` (of W in the grid) owns the slices wid, wid + W, wid + 2W,
// ... (up to MAXS per round), spread uniformly over the matrix: nearby rows put
// their entries in similar panels, so interleaving gives every workgroup about
// the average work per panel and the grid sweeps the panels in step.
template <bool REAL, bool LZ, int WPB, int MAXS>
__global__ void __launch_bounds__(WPB * 64) k_sps(PanelView P, const double2* __restrict__ x,
                                                  double2* __restrict__ out, const double2* __restrict__ Wp, Scal sc,
                                                  int j, double* partials) {
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t n = P.n;
    const int64_t nslices = (n + 63) >> 6;
    const int64_t W = (int64_t)gridDim.x * WPB;
    const int64_t wid = (int64_t)blockIdx.x * WPB + wv;
    double ib = 0, cp = 0;
    if (LZ) {
        ib = sc.invb[j];
        cp = j > 0 ? sc.beta[j] * sc.invb[j - 1] : 0.0;
    }
    __shared__ double2 yl[WPB * MAXS * 64];  // row sums of this wavefront's slices
    double2* my = yl + wv * MAXS * 64;
    double dacc = 0.0;
    for (int64_t r0 = 0; r0 < nslices; r0 += W * MAXS) {
        for (int k = 0; k < MAXS; ++k) my[k * 64 + lane] = make_double2(0.0, 0.0);
        for (int p = 0; p < P.npan; ++p) {
            __syncthreads();  // keep the workgroup on one panel
            const uint32_t* rp = P.prow + (int64_t)p * (n + 1);
            const int64_t base = P.pbase[p];
            const double2* xp = x + (int64_t)p * kPanelWidth;
#pragma unroll 1
            for (int k = 0; k < MAXS; ++k) {
                const int64_t slice = r0 + wid + (int64_t)k * W;
                if (slice >= nslices) continue;  // wave-uniform
                const int64_t row0 = slice * 64, myrow = row0 + lane;
                const int64_t rend = row0 + 64 < n ? row0 + 64 : n;
                const uint32_t s = myrow < n ? rp[myrow] : rp[rend];
                const uint32_t e = myrow < n ? rp[myrow + 1] : rp[rend];
                const uint32_t E0 = __shfl(s, 0, 64);
                const uint32_t E1 = __shfl(e, 63, 64);
                double yr = 0.0, yi = 0.0;
                for (uint32_t K = E0; K < E1; K += 64) {
                    const uint32_t kk = K + lane;
                    double pr = 0.0, pi = 0.0;
                    if (kk < E1) {
                        const int c = __builtin_nontemporal_load(&P.pcol[base + kk]);
                        const double2 xv = xp[c];
                        if (REAL) {
                            const double v = ntload1(&P.pvr[base + kk]);
                            pr = v * xv.x;
                            pi = v * xv.y;
                        } else {
                            const double2 v = ntload(&P.pvc[base + kk]);
                            pr = v.x * xv.x - v.y * xv.y;
                            pi = v.x * xv.y + v.y * xv.x;
                        }
                    }
                    const double Pr = wave_incl_scan(pr, lane), Pi = wave_incl_scan(pi, lane);
                    // this row's lanes in the window: [a, b]
                    const int a = (int)(s > K ? s - K : 0);
                    const int b = (int)((e < K + 64 ? e : K + 64) - K) - 1;
                    const bool mine = s < e && s < K + 64 && e > K;
                    const int bl = mine ? b : 0, al = mine && a > 0 ? a - 1 : 0;
                    const double hr = __shfl(Pr, bl, 64), hi = __shfl(Pi, bl, 64);
                    const double lr = __shfl(Pr, al, 64), li = __shfl(Pi, al, 64);
                    if (mine) {
                        yr += a > 0 ? hr - lr : hr;
                        yi += a > 0 ? hi - li : hi;
                    }
                }
                double2 t = my[k * 64 + lane];
                t.x += yr;
                t.y += yi;
                my[k * 64 + lane] = t;
            }
        }
        for (int k = 0; k < MAXS; ++k) {
            const int64_t slice = r0 + wid + (int64_t)k * W;
            const int64_t myrow = slice * 64 + lane;
            if (slice >= nslices || myrow >= n) continue;
            const double2 yv = my[k * 64 + lane];
            if (LZ) {
                double ur = ib * yv.x, ui = ib * yv.y;
                if (j > 0) {
                    const double2 q = Wp[myrow];
                    ur -= cp * q.x;
                    ui -= cp * q.y;
                }
                out[myrow] = make_double2(ur, ui);
                const double2 w = x[myrow];
                dacc += w.x * ur + w.y * ui;
            } else {
                out[myrow] = yv;
            }
        }
    }
    if (LZ) block_partial(dacc, partials);
}

// ---- jagged panel SpMV --------------------------------------------------------------
// Panels in step-major order (Panels::jagged, read by jag_row_sum): lane l
// owns row l of a 64-row slice and walks its own entries, every matrix load
// coalesced, each row summed in column order in registers, no value crossing
// lanes.  Slice ownership and the panel sweep are those of k_sps.
template <bool REAL, bool LZ, int WPB, int MAXS>
__global__ void __launch_bounds__(WPB * 64) k_jag(PanelView P, const double2* __restrict__ x,
                                                  double2* __restrict__ out, const double2* __restrict__ Wp, Scal sc,
                                                  int j, double* partials) {
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t n = P.n;
    const int64_t nslices = (n + 63) >> 6;
    const int64_t W = (int64_t)gridDim.x * WPB;
    const int64_t wid = (int64_t)blockIdx.x * WPB + wv;
    double ib = 0, cp = 0;
    if (LZ) {
        ib = sc.invb[j];
        cp = j > 0 ? sc.beta[j] * sc.invb[j - 1] : 0.0;
    }
    __shared__ double2 yl[WPB * MAXS * 64];  // row sums of this wavefront's slices
    double2* my = yl + wv * MAXS * 64;
    double dacc = 0.0;
    // Wavefront wid owns slices wid, wid + W, ... (T of them at most), taken
    // in R rounds of m <= MAXS each: the rounds carry equal shares, so no
    // round sweeps the panels with a single slice per wavefront.
    const int64_t T = (nslices + W - 1) / W;
    const int64_t R = (T + MAXS - 1) / MAXS;
    const int m = (int)((T + R - 1) / R);
    for (int64_t r0 = 0; r0 < R * m; r0 += m) {
        for (int k = 0; k < MAXS; ++k) my[k * 64 + lane] = make_double2(0.0, 0.0);
        for (int p = 0; p < P.npan; ++p) {
            __syncthreads();  // keep the workgroup on one panel
            const uint32_t* rp = P.prow + (int64_t)p * (n + 1);
            const int64_t base = P.pbase[p];
            const double2* xp = x + (int64_t)p * kPanelWidth;
#pragma unroll 1
            for (int k = 0; k < m; ++k) {
                const int64_t slice = wid + (r0 + k) * W;
                if (slice >= nslices) continue;  // wave-uniform
                const int64_t row0 = slice * 64, myrow = row0 + lane;
                const int64_t rend = row0 + 64 < n ? row0 + 64 : n;
                const uint32_t s = myrow < n ? rp[myrow] : rp[rend];
                const uint32_t e = myrow < n ? rp[myrow + 1] : rp[rend];
                const uint32_t len = e - s;
                const int64_t run = base + (uint32_t)__builtin_amdgcn_readfirstlane((int)s);  // lane 0: rp[row0]
                double ar = 0.0, ai = 0.0;
                jag_row_sum<REAL>(P.pcol + run, P.pvr + run, P.pvc + run, xp, len, ar, ai);
                double2 tacc = my[k * 64 + lane];
                tacc.x += ar;
                tacc.y += ai;
                my[k * 64 + lane] = tacc;
            }
        }
        for (int k = 0; k < m; ++k) {
            const int64_t slice = wid + (r0 + k) * W;
            const int64_t myrow = slice * 64 + lane;
            if (slice >= nslices || myrow >= n) continue;
            const double2 yv = my[k * 64 + lane];
            if (LZ) {
                double ur = ib * yv.x, ui = ib * yv.y;
                if (j > 0) {
                    const double2 q = Wp[myrow];
                    ur -= cp * q.x;
                    ui -= cp * q.y;
                }
                out[myrow] = make_double2(ur, ui);
                const double2 w = x[myrow];
                dacc += w.x * ur + w.y * ui;
            } else {
                out[myrow] = yv;
            }
        }
    }
    if (LZ) block_partial(dacc, partials);
}

// SpMV kind 5: as k_jag, but lane l of a wave sums rows l, l+64, ..,
// l+64(R-1) of a 64R-row super-slice in one stream per panel (panels laid out
// with R rows per lane), so the lanes' stream lengths even out and fewer lane
// slots idle per step.  MAXS super-slices per wave; R * MAXS <= 8.
template <bool REAL, bool LZ, int WPB, int R, int MAXS>
__global__ void __launch_bounds__(WPB * 64) k_jagr(PanelView P, const double2* __restrict__ x,
                                                   double2* __restrict__ out, const double2* __restrict__ Wp, Scal sc,
                                                   int j, double* partials) {
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t n = P.n;
    const int64_t nss = (n + 64 * R - 1) / (64 * R);
    const int64_t W = (int64_t)gridDim.x * WPB;
    const int64_t wid = (int64_t)blockIdx.x * WPB + wv;
    double ib = 0, cp = 0;
    if (LZ) {
        ib = sc.invb[j];
        cp = j > 0 ? sc.beta[j] * sc.invb[j - 1] : 0.0;
    }
    __shared__ double2 yl[WPB * MAXS * R * 64];  // row sums of this wavefront's super-slices
    double2* my = yl + wv * MAXS * R * 64;
    double dacc = 0.0;
    for (int64_t r0 = 0; r0 < nss; r0 += W * MAXS) {
        for (int k = 0; k < MAXS * R; ++k) my[k * 64 + lane] = make_double2(0.0, 0.0);
        for (int p = 0; p < P.npan; ++p) {
            __syncthreads();  // keep the workgroup on one panel
            const uint32_t* rp = P.prow + (int64_t)p * (n + 1);
            const int64_t base = P.pbase[p];
            const double2* xp = x + (int64_t)p * kPanelWidth;
#pragma unroll 1
            for (int k = 0; k < MAXS; ++k) {
                const int64_t ss = r0 + wid + (int64_t)k * W;
                if (ss >= nss) continue;  // wave-uniform
                const int64_t row0 = ss * 64 * R;
                uint32_t cum[R];
                uint32_t c = 0;
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    const int64_t row = row0 + lane + 64 * q;
                    c += row < n ? rp[row + 1] - rp[row] : 0u;
                    cum[q] = c;
                }
                const int64_t run = base + rp[row0];
                jag_rows_sum<REAL, R>(P.pcol + run, P.pvr + run, P.pvc + run, xp, cum, my + k * R * 64 + lane);
            }
        }
        for (int k = 0; k < MAXS; ++k)
            for (int q = 0; q < R; ++q) {
                const int64_t ss = r0 + wid + (int64_t)k * W;
                const int64_t myrow = ss * 64 * R + lane + 64 * q;
                if (ss >= nss || myrow >= n) continue;
                const double2 yv = my[(k * R + q) * 64 + lane];
                if (LZ) {
                    double ur = ib * yv.x, ui = ib * yv.y;
                    if (j > 0) {
                        const double2 w = Wp[myrow];
                        ur -= cp * w.x;
                        ui -= cp * w.y;
                    }
                    out[myrow] = make_double2(ur, ui);
                    const double2 w = x[myrow];
                    dacc += w.x * ur + w.y * ui;
                } else {
                    out[myrow] = yv;
                }
            }
    }
    if (LZ) block_partial(dacc, partials);
}

// ---- column-tile SpMV (kind 4) ---------------------------------------------------------
// Workgroup b (16 wavefronts) owns a row block; lane l of wave w keeps the
// sums of its R rows (one per owned slice) in registers for the whole
// product.  The columns are swept in tiles of kTileCols: the x tile is staged
// in LDS (the next tile is loaded into registers while the current one is
// consumed).  Within a tile the wave's entries are stored step-major across
// its R rows (tiles.hip), so at step j the wave issues the loads of all R
// rows before using any: enough bytes in flight per wave to stream HBM, and
// every gather is an LDS read.  HBM traffic: the matrix stream, the u8 entry
// counts, x once per XCD and the vectors of the epilogue.
constexpr int kT4Threads = kTileWaves * 64;
constexpr int kT4Pre = kTileCols / kT4Threads;  // x entries per thread per tile

template <bool REAL, bool LZ, int R>
__global__ void __launch_bounds__(kT4Threads) k_t4(TileView V, const double2* __restrict__ x,
                                                   double2* __restrict__ out, const double2* __restrict__ Wp, Scal sc,
                                                   int j, double* partials) {
    __shared__ double2 xs[kTileCols];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t n = V.n;
    const int64_t nslices = (n + 63) >> 6;
    const int64_t wid = (int64_t)blockIdx.x * kTileWaves + wv;
    const int64_t s0 = (int64_t)blockIdx.x * V.SB;
    const int64_t s1 = s0 + V.SB < nslices ? s0 + V.SB : nslices;
    double ar[R], ai[R];
#pragma unroll
    for (int r = 0; r < R; ++r) ar[r] = ai[r] = 0.0;
    int64_t ptr = V.wstart[wid];
    const uint8_t* L = V.len + wid * (int64_t)V.ntiles * V.R * 64 + lane;
    double2 pre[kT4Pre];
    auto load_tile = [&](int t) {
#pragma unroll
        for (int k = 0; k < kT4Pre; ++k) {
            const int64_t c = (int64_t)t * kTileCols + k * kT4Threads + threadIdx.x;
            pre[k] = c < n ? x[c] : make_double2(0.0, 0.0);
        }
    };
    load_tile(0);
    for (int t = 0; t < V.ntiles; ++t) {
        __syncthreads();  // every wave is done with the previous tile
#pragma unroll
        for (int k = 0; k < kT4Pre; ++k) xs[k * kT4Threads + threadIdx.x] = pre[k];
        __syncthreads();
        if (t + 1 < V.ntiles) load_tile(t + 1);  // in flight while this tile is consumed
        uint32_t l[R];
        uint32_t lmax = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            l[r] = r < V.R ? (uint32_t)L[((int64_t)t * V.R + r) * 64] : 0u;
            lmax = l[r] > lmax ? l[r] : lmax;
        }
        const uint16_t* cr = V.col + ptr;
        const double* vr = V.vr + ptr;
        const double2* vc = V.vc + ptr;
        uint32_t off = 0;
        for (uint32_t st = 0; __ballot(st < lmax) != 0; ++st) {
            uint32_t q[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const uint64_t m = __ballot(st < l[r]);
                q[r] = off + lanes_below(m);
                off += (uint32_t)__popcll(m);
            }
            int c[R];
            double2 v[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const bool act = st < l[r];
                c[r] = act ? (int)__builtin_nontemporal_load(&cr[q[r]]) : 0;
                if (REAL) v[r] = make_double2(act ? ntload1(&vr[q[r]]) : 0.0, 0.0);
                else v[r] = act ? ntload(&vc[q[r]]) : make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const double2 xv = xs[c[r]];
                if (REAL) {
                    ar[r] += v[r].x * xv.x;
                    ai[r] += v[r].x * xv.y;
                } else {
                    ar[r] += v[r].x * xv.x - v[r].y * xv.y;
                    ai[r] += v[r].x * xv.y + v[r].y * xv.x;
                }
            }
        }
        ptr += off;
    }
    double ib = 0, cp = 0;
    if (LZ) {
        ib = sc.invb[j];
        cp = j > 0 ? sc.beta[j] * sc.invb[j - 1] : 0.0;
    }
    double dacc = 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int64_t slice = s0 + wv + (int64_t)kTileWaves * r;
        const int64_t row = slice * 64 + lane;
        if (r < V.R && slice < s1 && row < n) {
            if (LZ) {
                double ur = ib * ar[r], ui = ib * ai[r];
                if (j > 0) {
                    const double2 q = Wp[row];
                    ur -= cp * q.x;
                    ui -= cp * q.y;
                }
                out[row] = make_double2(ur, ui);
                const double2 w = x[row];
                dacc += w.x * ur + w.y * ui;
            } else {
                out[row] = make_double2(ar[r], ai[r]);
            }
        }
    }
    if (LZ) block_partial(dacc, partials);
}

// ---- launch helpers -----------------------------------------------------------
int vec_grid(int64_t n) {
    const int ncu = runtime().num_cus;
    return (int)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, kThreads), (int64_t)ncu * 8));
}

int spmv_grid(int64_t n, int lpr) {
    const int ncu = runtime().num_cus;
    const int64_t rpb = kThreads / lpr;
    return (int)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, rpb), (int64_t)ncu * 16));
}

#define BLQH_LPR_SWITCH(lpr, KERNEL, ...)                      \
    switch (lpr) {                                             \
        case 4: hipLaunchKernelGGL(KERNEL<4>, __VA_ARGS__); break;   \
        case 8: hipLaunchKernelGGL(KERNEL<8>, __VA_ARGS__); break;   \
        case 16: hipLaunchKernelGGL(KERNEL<16>, __VA_ARGS__); break; \
        case 32: hipLaunchKernelGGL(KERNEL<32>, __VA_ARGS__); break; \
        default: hipLaunchKernelGGL(KERNEL<64>, __VA_ARGS__); break; \
    }

// SpMV variant: row-major CSR with `lpr` lanes per row, or the panel kernel
// with `lpr` lanes per row segment, `threads` per workgroup and `wpc`
// persistent workgroups per CU.
struct Plan {
    int kind, lpr, threads, wpc, rows;  // rows: rows per lane group in flight (panel kernel)
};

struct PanelLaunch {
    int grid;
    int64_t rows_per_wg;
    int chunk;
    size_t smem;
};

PanelLaunch panel_launch(const blqh_hamil* H, const Plan& pl) {
    const int ncu = runtime().num_cus;
    PanelLaunch L;
    L.grid = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)ncu * pl.wpc, ceil_div(H->n, 64)));
    L.rows_per_wg = ceil_div(H->n, L.grid);
    const int64_t lds_rows = (pl.wpc == 1 ? 144 * 1024 : 72 * 1024) / (int64_t)sizeof(double2);
    L.chunk = (int)std::max<int64_t>(1, std::min<int64_t>(L.rows_per_wg, lds_rows));
    L.smem = (size_t)L.chunk * sizeof(double2);
    return L;
}

template <int LPR, int PT, int U, bool REAL, bool LZ>
void panel_go(const blqh_hamil* H, const PanelLaunch& L, const double2* x, double2* out, const double2* Wp,
              const Scal& sc, int j, double* partials, hipStream_t st) {
    auto kern = k_pspmv<LPR, PT, U, REAL, LZ>;
    static const bool attr = [&] {  // once per instantiation, thread-safe
        BLQH_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
        return true;
    }();
    (void)attr;
    hipLaunchKernelGGL(kern, dim3(L.grid), dim3(PT), L.smem, st, view(H->pan, H->n), x, out, L.rows_per_wg, L.chunk,
                       Wp, sc, j, partials);
}

template <bool LZ>
void panel_dispatch(const blqh_hamil* H, const Plan& pl, const double2* x, double2* out, const double2* Wp,
                    const Scal& sc, int j, double* partials, hipStream_t st) {
    BLQH_REQUIRE(!H->pan.jagged, BLQH_EINTERNAL, "row-segment SpMV on a jagged panel layout");
    const PanelLaunch L = panel_launch(H, pl);
    const bool real = H->pan.real;
#define BLQH_PANEL_CASE(LPR, PT, U)                                                                  \
    if (pl.lpr == LPR && pl.threads == PT && pl.rows == U) {                                          \
        if (real) panel_go<LPR, PT, U, true, LZ>(H, L, x, out, Wp, sc, j, partials, st);     \
        else panel_go<LPR, PT, U, false, LZ>(H, L, x, out, Wp, sc, j, partials, st);         \
        return;                                                                                       \
    }
    BLQH_PANEL_CASE(4, 512, 2) BLQH_PANEL_CASE(8, 512, 2) BLQH_PANEL_CASE(16, 512, 2)
    BLQH_PANEL_CASE(4, 1024, 2) BLQH_PANEL_CASE(8, 1024, 2) BLQH_PANEL_CASE(16, 1024, 2)
    BLQH_PANEL_CASE(4, 512, 4) BLQH_PANEL_CASE(8, 512, 4) BLQH_PANEL_CASE(16, 512, 4)
    BLQH_PANEL_CASE(4, 1024, 4) BLQH_PANEL_CASE(8, 1024, 4) BLQH_PANEL_CASE(16, 1024, 4)
#undef BLQH_PANEL_CASE
    throw Error(BLQH_EINTERNAL, "unsupported panel SpMV plan");
}

Plan plan_of(const blqh_hamil* H) {
    return Plan{H->spmv_kind, H->spmv_lpr, H->spmv_threads, H->spmv_wgs_per_cu, H->spmv_rows};
}

int sps_grid(const blqh_hamil* H, const Plan& pl) {
    const int64_t groups = ceil_div(ceil_div(H->n, 64), (int64_t)(pl.threads / 64) * pl.rows);
    return (int)std::max<int64_t>(1, std::min<int64_t>(groups, (int64_t)runtime().num_cus * pl.wpc));
}

template <bool LZ>
void sps_dispatch(const blqh_hamil* H, const Plan& pl, const double2* x, double2* out, const double2* Wp,
                  const Scal& sc, int j, double* partials, hipStream_t st) {
    BLQH_REQUIRE(!H->pan.jagged, BLQH_EINTERNAL, "slice-streaming SpMV on a jagged panel layout");
    const PanelView V = view(H->pan, H->n);
    const dim3 grid(sps_grid(H, pl)), block(pl.threads);
#define BLQH_SPS_CASE(WPB, MAXS)                                                                            \
    if (pl.threads == WPB * 64 && pl.rows == MAXS) {                                                        \
        if (H->pan.real)                                                                                    \
            hipLaunchKernelGGL((k_sps<true, LZ, WPB, MAXS>), grid, block, 0, st, V, x, out, Wp, sc, j, partials); \
        else                                                                                                \
            hipLaunchKernelGGL((k_sps<false, LZ, WPB, MAXS>), grid, block, 0, st, V, x, out, Wp, sc, j, partials); \
        return;                                                                                             \
    }
    BLQH_SPS_CASE(8, 1) BLQH_SPS_CASE(8, 2) BLQH_SPS_CASE(8, 4) BLQH_SPS_CASE(8, 8)
    BLQH_SPS_CASE(16, 1) BLQH_SPS_CASE(16, 2) BLQH_SPS_CASE(16, 4) BLQH_SPS_CASE(16, 8)
#undef BLQH_SPS_CASE
    throw Error(BLQH_EINTERNAL, "unsupported slice-streaming SpMV plan");
}

template <bool LZ>
void jag_dispatch(const blqh_hamil* H, const Plan& pl, const double2* x, double2* out, const double2* Wp,
                  const Scal& sc, int j, double* partials, hipStream_t st) {
    BLQH_REQUIRE(H->pan.jagged && H->pan.rows == 1, BLQH_EINTERNAL, "jagged SpMV without its panel layout");
    const PanelView V = view(H->pan, H->n);
    const dim3 grid(sps_grid(H, pl)), block(pl.threads);
#define BLQH_JAG_CASE(WPB, MAXS)                                                                            \
    if (pl.threads == WPB * 64 && pl.rows == MAXS) {                                                        \
        if (H->pan.real)                                                                                    \
            hipLaunchKernelGGL((k_jag<true, LZ, WPB, MAXS>), grid, block, 0, st, V, x, out, Wp, sc, j, partials); \
        else                                                                                                \
            hipLaunchKernelGGL((k_jag<false, LZ, WPB, MAXS>), grid, block, 0, st, V, x, out, Wp, sc, j, partials); \
        return;                                                                                             \
    }
    BLQH_JAG_CASE(4, 1) BLQH_JAG_CASE(4, 2) BLQH_JAG_CASE(4, 4) BLQH_JAG_CASE(4, 8)
    BLQH_JAG_CASE(8, 1) BLQH_JAG_CASE(8, 2) BLQH_JAG_CASE(8, 4) BLQH_JAG_CASE(8, 8)
    BLQH_JAG_CASE(16, 1) BLQH_JAG_CASE(16, 2) BLQH_JAG_CASE(16, 4) BLQH_JAG_CASE(16, 8)
#undef BLQH_JAG_CASE
    throw Error(BLQH_EINTERNAL, "unsupported jagged SpMV plan");
}

int jagr_grid(const blqh_hamil* H, const Plan& pl) {
    const int64_t nss = ceil_div(H->n, (int64_t)64 * pl.lpr);
    const int64_t groups = ceil_div(nss, (int64_t)(pl.threads / 64) * pl.rows);
    return (int)std::max<int64_t>(1, std::min<int64_t>(groups, (int64_t)runtime().num_cus * pl.wpc));
}

template <bool LZ>
void jagr_dispatch(const blqh_hamil* H, const Plan& pl, const double2* x, double2* out, const double2* Wp,
                   const Scal& sc, int j, double* partials, hipStream_t st) {
    BLQH_REQUIRE(H->pan.jagged && H->pan.rows == pl.lpr, BLQH_EINTERNAL,
                 "multi-row jagged SpMV without its panel layout");
    const PanelView V = view(H->pan, H->n);
    const dim3 grid(jagr_grid(H, pl)), block(pl.threads);
#define BLQH_JAGR_CASE(WPB, R, MAXS)                                                                            \
    if (pl.threads == WPB * 64 && pl.lpr == R && pl.rows == MAXS) {                                             \
        if (H->pan.real)                                                                                        \
            hipLaunchKernelGGL((k_jagr<true, LZ, WPB, R, MAXS>), grid, block, 0, st, V, x, out, Wp, sc, j, partials); \
        else                                                                                                    \
            hipLaunchKernelGGL((k_jagr<false, LZ, WPB, R, MAXS>), grid, block, 0, st, V, x, out, Wp, sc, j, partials); \
        return;                                                                                                 \
    }
    BLQH_JAGR_CASE(8, 2, 1) BLQH_JAGR_CASE(8, 2, 2) BLQH_JAGR_CASE(8, 2, 4)
    BLQH_JAGR_CASE(8, 4, 1) BLQH_JAGR_CASE(8, 4, 2) BLQH_JAGR_CASE(8, 8, 1)
    BLQH_JAGR_CASE(16, 2, 1) BLQH_JAGR_CASE(16, 2, 2) BLQH_JAGR_CASE(16, 2, 4)
    BLQH_JAGR_CASE(16, 4, 1) BLQH_JAGR_CASE(16, 4, 2) BLQH_JAGR_CASE(16, 8, 1)
#undef BLQH_JAGR_CASE
    throw Error(BLQH_EINTERNAL, "unsupported multi-row jagged SpMV plan");
}

// smallest compiled slice count per wave that covers the layout's R
int t4_R(int R) {
    for (int c : {1, 2, 3, 4, 6, 8, 9, 10, 12, 16})
        if (c >= R) return c;
    return -1;
}

template <bool LZ>
void t4_dispatch(const blqh_hamil* H, const double2* x, double2* out, const double2* Wp, const Scal& sc, int j,
                 double* partials, hipStream_t st) {
    const Tiles& T = H->til;
    BLQH_REQUIRE(T.built, BLQH_EINTERNAL, "column-tile SpMV without its layout");
    const TileView V = tview(T, H->n);
    const dim3 grid(T.G), block(kT4Threads);
#define BLQH_T4_CASE(RR)                                                                                    \
    if (t4_R(T.R) == RR) {                                                                                  \
        if (T.real)                                                                                         \
            hipLaunchKernelGGL((k_t4<true, LZ, RR>), grid, block, 0, st, V, x, out, Wp, sc, j, partials); \
        else                                                                                                \
            hipLaunchKernelGGL((k_t4<false, LZ, RR>), grid, block, 0, st, V, x, out, Wp, sc, j, partials); \
        return;                                                                                             \
    }
    BLQH_T4_CASE(1) BLQH_T4_CASE(2) BLQH_T4_CASE(3) BLQH_T4_CASE(4) BLQH_T4_CASE(6) BLQH_T4_CASE(8)
    BLQH_T4_CASE(9) BLQH_T4_CASE(10) BLQH_T4_CASE(12) BLQH_T4_CASE(16)
#undef BLQH_T4_CASE
    throw Error(BLQH_EINTERNAL, "unsupported column-tile SpMV layout");
}

void launch_spmv_plan(const blqh_hamil* H, const Plan& pl, const double2* x, double2* y, hipStream_t st) {
    if (pl.kind == 5) {
        jagr_dispatch<false>(H, pl, x, y, nullptr, Scal{}, 0, nullptr, st);
    } else if (pl.kind == 4) {
        t4_dispatch<false>(H, x, y, nullptr, Scal{}, 0, nullptr, st);
    } else if (pl.kind == 3) {
        jag_dispatch<false>(H, pl, x, y, nullptr, Scal{}, 0, nullptr, st);
    } else if (pl.kind == 2) {
        sps_dispatch<false>(H, pl, x, y, nullptr, Scal{}, 0, nullptr, st);
    } else if (pl.kind == 1) {
        panel_dispatch<false>(H, pl, x, y, nullptr, Scal{}, 0, nullptr, st);
    } else {
        const CsrView A = spmv_csr(H);
        BLQH_LPR_SWITCH(pl.lpr, k_spmv, dim3(spmv_grid(H->n, pl.lpr)), dim3(kThreads), 0, st, H->n, A.rowptr,
                        A.col, A.val, x, y);
    }
}

// grid size (for partial-sum buffers) of the fused K1 under a plan
int lz1_grid(const blqh_hamil* H, const Plan& pl) {
    if (pl.kind == 4) return H->til.G;
    if (pl.kind == 5) return jagr_grid(H, pl);
    return pl.kind >= 2 ? sps_grid(H, pl) : pl.kind == 1 ? panel_launch(H, pl).grid : spmv_grid(H->n, pl.lpr);
}

void launch_lz1(const blqh_hamil* H, const Plan& pl, const double2* W, const double2* Wp, double2* u, const Scal& sc,
                int j, double* partials, hipStream_t st) {
    if (pl.kind == 5) {
        jagr_dispatch<true>(H, pl, W, u, Wp, sc, j, partials, st);
    } else if (pl.kind == 4) {
        t4_dispatch<true>(H, W, u, Wp, sc, j, partials, st);
    } else if (pl.kind == 3) {
        jag_dispatch<true>(H, pl, W, u, Wp, sc, j, partials, st);
    } else if (pl.kind == 2) {
        sps_dispatch<true>(H, pl, W, u, Wp, sc, j, partials, st);
    } else if (pl.kind == 1) {
        panel_dispatch<true>(H, pl, W, u, Wp, sc, j, partials, st);
    } else {
        const CsrView A = spmv_csr(H);
        BLQH_LPR_SWITCH(pl.lpr, k_lz1, dim3(spmv_grid(H->n, pl.lpr)), dim3(kThreads), 0, st, H->n, A.rowptr,
                        A.col, A.val, W, Wp, u, sc, j, partials);
    }
}

// Shape of a tuned matrix (the plan cache's key).
struct TuneKey {
    int npan, real, lg_n, lg_row;  // log2 of rows and of the mean row length, in quarters
    bool operator==(const TuneKey& o) const {
        return npan == o.npan && real == o.real && lg_n == o.lg_n && lg_row == o.lg_row;
    }
};
std::mutex g_tune_mu;
std::vector<std::pair<TuneKey, Plan>> g_tune_cache;  // most recent last, at most 64

// Autotuner: time every variant on this matrix (random x) and keep the fastest.
void tune(blqh_hamil* H) {
    // BLQH_SPMV_PLAN="kind,lanes,threads,wgs_per_cu,rows" pins a variant (profiling, reproducibility)
    if (const char* env = std::getenv("BLQH_SPMV_PLAN")) {
        Plan p{};
        if (std::sscanf(env, "%d,%d,%d,%d,%d", &p.kind, &p.lpr, &p.threads, &p.wpc, &p.rows) == 5 &&
            (p.kind == 0 || H->pan.built) && (p.kind != 4 || build_tiles(H, runtime().num_cus))) {
            H->spmv_kind = p.kind;
            H->spmv_lpr = p.lpr;
            H->spmv_threads = p.threads;
            H->spmv_wgs_per_cu = p.wpc;
            H->spmv_rows = p.rows;
            return;
        }
    }
    // A matrix of the shape of one tuned before (same panels and value kind,
    // rows and mean row length within 19 %) takes that plan without timing:
    // the sectors of a sweep tune once.  BLQH_TUNE_CACHE=0 times every matrix.
    const double mean_row = H->n ? (double)H->nnz / H->n : 0;
    const char* cenv = std::getenv("BLQH_TUNE_CACHE");
    const bool use_cache = !(cenv && cenv[0] == '0');
    const TuneKey key{H->pan.built ? H->pan.npan : 0, H->pan.real ? 1 : 0,
                      (int)std::lround(4.0 * std::log2((double)std::max<int64_t>(H->n, 1))),
                      (int)std::lround(4.0 * std::log2(std::max(mean_row, 1.0)))};
    if (use_cache) {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        for (const auto& e : g_tune_cache)
            if (e.first == key && (e.second.kind == 0 || H->pan.built)) {
                const Plan& p = e.second;
                H->spmv_kind = p.kind;
                H->spmv_lpr = p.lpr;
                H->spmv_threads = p.threads;
                H->spmv_wgs_per_cu = p.wpc;
                H->spmv_rows = p.rows;
                return;
            }
    }
    hipStream_t st = stream();
    std::vector<Plan> cands;
    for (int lpr : {4, 8, 16, 32, 64})
        if (!(lpr > 4 * mean_row + 8 && lpr > 8)) cands.push_back(Plan{0, lpr, kThreads, 0, 0});
    // Kinds 1, 2, 4 and 5 never won on any measured matrix (79 to 2.35M rows);
    // BLQH_TUNE_ALL=1 times them too.
    const char* aenv = std::getenv("BLQH_TUNE_ALL");
    const bool all_kinds = aenv && aenv[0] == '1';
    if (H->pan.built && all_kinds)
        for (int pt : {512, 1024})
            for (int lpr : {4, 8, 16})
                for (int wpc : {1, 2})
                    for (int u : {2, 4}) cands.push_back(Plan{1, lpr, pt, wpc, u});
    if (H->pan.built && all_kinds)
        for (int pt : {512, 1024})
            for (int wpc : {1, 2, 4}) {
                // slices per wavefront: the smallest supported count that covers the matrix in one round
                const int64_t waves = (int64_t)runtime().num_cus * wpc * (pt / 64);
                const int64_t need = ceil_div(ceil_div(H->n, 64), waves);
                int ms = 1;
                while (ms < need && ms < 8) ms <<= 1;
                cands.push_back(Plan{2, 64, pt, wpc, ms});
            }
    // column tiles: an x tile per LDS, worth it when x does not fit the caches
    if (H->pan.built && all_kinds && H->n > 4 * kTileCols && build_tiles(H, runtime().num_cus))
        cands.push_back(Plan{4, 64, kT4Threads, 1, H->til.R});
    if (H->pan.built)  // jagged kernel last: one layout conversion
        for (int pt : {256, 512, 1024})
            for (int wpc : {1, 2, 4}) {
                const int64_t waves = (int64_t)runtime().num_cus * wpc * (pt / 64);
                const int64_t need = ceil_div(ceil_div(H->n, 64), waves);
                int ms = 1;
                while (ms < need && ms < 8) ms <<= 1;
                cands.push_back(Plan{3, 64, pt, wpc, ms});
            }
    if (H->pan.built && all_kinds)  // R rows per lane: one layout conversion per R
        for (int R : {2, 4, 8})
            for (int pt : {512, 1024})
                for (int wpc : {1, 2}) {
                    const int64_t waves = (int64_t)runtime().num_cus * wpc * (pt / 64);
                    const int64_t need = ceil_div(ceil_div(H->n, (int64_t)64 * R), waves);
                    int ms = 1;
                    while (ms < need && ms * R < 8) ms <<= 1;
                    cands.push_back(Plan{5, R, pt, wpc, ms});
                }
    std::vector<double2> hx(H->n);
    std::mt19937 rng(12345);
    for (auto& z : hx) z = make_double2(rng() * 1.0 / rng.max() - 0.5, rng() * 1.0 / rng.max() - 0.5);
    DevBuf<double2> x(H->n), y(H->n);
    BLQH_HIP(hipMemcpyAsync(x.get(), hx.data(), H->n * sizeof(double2), hipMemcpyHostToDevice, st));
    hipEvent_t e0, e1;
    BLQH_HIP(hipEventCreate(&e0));
    BLQH_HIP(hipEventCreate(&e1));
    float best = 1e30f;
    Plan bp = cands.front();
    std::vector<std::pair<Plan, float>> timed;
    // repetitions: at least about 2 ms of work per variant, so that small
    // matrices (tens of microseconds per product) are not timed on launch noise
    int reps = H->nnz > 50000000 ? 3 : 10;
    {
        const Plan& p0 = cands.front();
        launch_spmv_plan(H, p0, x.get(), y.get(), st);
        BLQH_HIP(hipEventRecord(e0, st));
        launch_spmv_plan(H, p0, x.get(), y.get(), st);
        BLQH_HIP(hipEventRecord(e1, st));
        BLQH_HIP(hipEventSynchronize(e1));
        float probe = 0;
        BLQH_HIP(hipEventElapsedTime(&probe, e0, e1));
        reps = std::max(reps, std::min(200, (int)std::ceil(2.0f / std::max(probe, 1e-3f))));
    }
    const bool log = std::getenv("BLQH_TUNE_LOG") != nullptr;  // print the timing of every variant
    for (const Plan& pl : cands) {
        if (pl.kind >= 1) set_panel_layout(H, pl.kind == 3 || pl.kind == 5, nullptr, 0, pl.kind == 5 ? pl.lpr : 1);
        launch_spmv_plan(H, pl, x.get(), y.get(), st);  // warm
        BLQH_HIP(hipEventRecord(e0, st));
        for (int r = 0; r < reps; ++r) launch_spmv_plan(H, pl, x.get(), y.get(), st);
        BLQH_HIP(hipEventRecord(e1, st));
        BLQH_HIP(hipEventSynchronize(e1));
        float ms = 0;
        BLQH_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (log) std::fprintf(stderr, "blqh spmv tune: kind %d lanes %d threads %d wgs/cu %d rows %d: %.3f ms\n",
                              pl.kind, pl.lpr, pl.threads, pl.wpc, pl.rows, ms / reps);
        timed.push_back({pl, ms});
        if (ms < best) { best = ms; bp = pl; }
    }
    // Near-ties (within 2 %) of the winning panel kind go to the largest
    // workgroup, then the fewest workgroups per CU: the per-panel barrier keeps
    // a whole workgroup on one panel, so bigger workgroups hold a smaller x
    // working set per CU (2.35M-state N_phi = 16 sector: 1024 x 1 moves 13.5 GB per launch where
    // 512 x 2 moves 23 GB, in the same time).
    if (bp.kind >= 2)
        for (const auto& c : timed)
            if (c.first.kind == bp.kind && c.second <= best * 1.02f &&
                (c.first.threads > bp.threads || (c.first.threads == bp.threads && c.first.wpc < bp.wpc)))
                bp = c.first;
    BLQH_HIP(hipGetLastError());
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    H->spmv_kind = bp.kind;
    H->spmv_lpr = bp.lpr;
    H->spmv_threads = bp.threads;
    H->spmv_wgs_per_cu = bp.wpc;
    H->spmv_rows = bp.rows;
    if (use_cache) {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        if (g_tune_cache.size() >= 64) g_tune_cache.erase(g_tune_cache.begin());
        g_tune_cache.push_back({key, bp});
    }
}

// Tunes on first use (unless a variant is pinned) and puts the panels in the
// entry order the chosen kernel reads; the column-tile copy is released when
// another variant won.
void autotune(blqh_hamil* H) {
    ensure_panels(H);
    if (H->spmv_kind < 0) {
        tune(H);
        if (H->spmv_kind != 4 && H->til.built) {
            Tiles& T = H->til;
            T.len.release();
            T.wstart.release();
            T.col.release();
            T.vr.release();
            T.vc.release();
            T.built = false;
            T.tried = false;
        }
    }
    if ((H->spmv_kind >= 1 && H->spmv_kind <= 3) || H->spmv_kind == 5)
        set_panel_layout(H, H->spmv_kind == 3 || H->spmv_kind == 5, nullptr, 0, H->spmv_kind == 5 ? H->spmv_lpr : 1);
}

}  // namespace

// ---- host: start vector and tridiagonal eigensolver ------------------------------
namespace blqh {

namespace {
// The std::mt19937 sequence (Matsumoto & Nishimura's MT19937: 624-word
// state, the standard's default seeding and tempering) produced a whole
// state at a time: the twist and the tempering are plain loops over the
// block, several times faster than one call per draw; same words.
struct Mt19937Blocks {
    uint32_t mt[624];
    explicit Mt19937Blocks(uint32_t seed) {
        mt[0] = seed;
        for (uint32_t i = 1; i < 624; ++i) mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + i;
    }
    // the next 624 draws
    void next(uint32_t* out) {
        constexpr uint32_t kUp = 0x80000000u, kLo = 0x7fffffffu, kA = 0x9908b0dfu;
        int i = 0;
        for (; i < 227; ++i) {
            const uint32_t y = (mt[i] & kUp) | (mt[i + 1] & kLo);
            mt[i] = mt[i + 397] ^ (y >> 1) ^ ((y & 1u) ? kA : 0u);
        }
        for (; i < 623; ++i) {
            const uint32_t y = (mt[i] & kUp) | (mt[i + 1] & kLo);
            mt[i] = mt[i - 227] ^ (y >> 1) ^ ((y & 1u) ? kA : 0u);
        }
        const uint32_t y = (mt[623] & kUp) | (mt[0] & kLo);
        mt[623] = mt[396] ^ (y >> 1) ^ ((y & 1u) ? kA : 0u);
        for (int k = 0; k < 624; ++k) {
            uint32_t z = mt[k];
            z ^= z >> 11;
            z ^= (z << 7) & 0x9d2c5680u;
            z ^= (z << 15) & 0xefc60000u;
            z ^= z >> 18;
            out[k] = z;
        }
    }
};
}  // namespace

// The reference's start vector before normalisation: entry i takes its
// imaginary part from draw 2i and its real part from draw 2i+1 of
// std::mt19937(seed) (the evaluation order g++ gives to the two rng() calls
// in the reference's complex<double>(re, im) constructor), each
// draw/max - 0.5.  Returns the running sum of |z_i|^2 in entry order.
double start_entries(int64_t n, uint32_t seed, double2* v) {
    Mt19937Blocks rng(seed);
    uint32_t raw[624];
    int used = 624;
    double nrm = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (used == 624) {  // 624 is even: an entry's two draws share a block
            rng.next(raw);
            used = 0;
        }
        const double im = raw[used] * 1.0 / 4294967295u - 0.5;
        const double re = raw[used + 1] * 1.0 / 4294967295u - 0.5;
        used += 2;
        v[i] = make_double2(re, im);
        nrm += re * re + im * im;
    }
    return nrm;
}

// start_entries normalised by the square root of their running sum.
void start_vector(int64_t n, uint32_t seed, std::vector<double2>& v) {
    v.resize(n);
    const double nrm = std::sqrt(start_entries(n, seed, v.data()));
    for (int64_t i = 0; i < n; ++i) { v[i].x /= nrm; v[i].y /= nrm; }
}

// Number of eigenvalues of T below x (Sturm sequence of the LDL^T pivots).
static int sturm_count(int m, const double* a, const double* b, double x) {
    int c = 0;
    double q = 1.0;
    for (int i = 0; i < m; ++i) {
        const double bb = i ? b[i - 1] * b[i - 1] : 0.0;
        q = a[i] - x - (i ? bb / q : 0.0);
        if (q == 0.0) q = -1e-300;
        if (q < 0) ++c;
    }
    return c;
}

// Lowest eigenvalue by bisection, eigenvector by inverse iteration, with
// caller work space: ws holds 8 m doubles, piv m ints (both used only when
// vec is given).
double tridiag_lowest_ws(int m, const double* a, const double* b, double* vec, double* ws, int* piv) {
    double lo = a[0], hi = a[0];
    for (int i = 0; i < m; ++i) {
        const double r = (i ? std::fabs(b[i - 1]) : 0.0) + (i + 1 < m ? std::fabs(b[i]) : 0.0);
        lo = std::min(lo, a[i] - r);
        hi = std::max(hi, a[i] + r);
    }
    for (int it = 0; it < 200; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        if (sturm_count(m, a, b, mid) >= 1) hi = mid; else lo = mid;
    }
    const double e = 0.5 * (lo + hi);
    if (!vec) return e;
    // inverse iteration: (T - e I) x = y with partial pivoting (Gaussian elimination on the band)
    double* x = ws;
    double* y = ws + m;
    double* d0 = ws + 2 * m;
    double* d1 = ws + 3 * m;
    double* d2 = ws + 4 * m;
    double* rhs = ws + 5 * m;
    double* l = ws + 6 * m;
    double* sub = ws + 7 * m;
    for (int i = 0; i < m; ++i) x[i] = 1.0 / std::sqrt((double)m);
    const double shift = e - 1e-14 * std::max(1.0, std::fabs(e));
    for (int iter = 0; iter < 4; ++iter) {
        // band LU of (T - shift I) with row pivoting: rows hold (d0, d1, d2) = cols (i, i+1, i+2)
        for (int i = 0; i < m; ++i) {
            d0[i] = a[i] - shift;
            d1[i] = i + 1 < m ? b[i] : 0.0;
            d2[i] = 0.0;
            rhs[i] = x[i];
            l[i] = 0.0;
            piv[i] = 0;
            sub[i] = 0.0;
        }
        // sub-diagonal of row i+1 is b[i]
        for (int i = 0; i + 1 < m; ++i) sub[i + 1] = b[i];
        for (int i = 0; i + 1 < m; ++i) {
            // rows i (pivot candidate) and i+1 (entry sub[i+1] in column i)
            if (std::fabs(sub[i + 1]) > std::fabs(d0[i])) {
                // swap rows i and i+1: row i+1 has (sub, d0, d1) in cols (i, i+1, i+2)
                std::swap(d0[i], sub[i + 1]);
                std::swap(d1[i], d0[i + 1]);
                std::swap(d2[i], d1[i + 1]);
                std::swap(rhs[i], rhs[i + 1]);
                piv[i] = 1;
            }
            if (d0[i] == 0.0) d0[i] = 1e-300;
            const double f = sub[i + 1] / d0[i];
            l[i] = f;
            d0[i + 1] -= f * d1[i];
            d1[i + 1] -= f * d2[i];
            rhs[i + 1] -= f * rhs[i];
        }
        if (d0[m - 1] == 0.0) d0[m - 1] = 1e-300;
        for (int i = m - 1; i >= 0; --i) {
            double s = rhs[i];
            if (i + 1 < m) s -= d1[i] * y[i + 1];
            if (i + 2 < m) s -= d2[i] * y[i + 2];
            y[i] = s / d0[i];
        }
        double nr = 0;
        for (int i = 0; i < m; ++i) nr += y[i] * y[i];
        nr = std::sqrt(nr);
        for (int i = 0; i < m; ++i) x[i] = y[i] / nr;
    }
    // sign convention: first component non-negative
    const double sg = x[0] < 0 ? -1.0 : 1.0;
    for (int i = 0; i < m; ++i) vec[i] = sg * x[i];
    return e;
}

// The lowest eigenvalue of T_m when `prev`, that of T_{m-1}, and `step`, the
// last change of the sequence, are known (the Ritz values of pass 1): by
// interlacing the eigenvalue lies at or below prev, so the bisection starts
// from [prev - w, prev + 1e-13 |prev|] with w = 4 step, widened 16-fold while
// the lower end is not below the eigenvalue (checked by Sturm counts), down to
// the Gershgorin bound.  About 25 bisection steps instead of 52 once the
// sequence settles.  Falls back to the full bracket without a usable prev.
double tridiag_lowest_next(int m, const double* a, const double* b, double prev, double step) {
    if (m < 3 || !std::isfinite(prev) || !std::isfinite(step))
        return tridiag_lowest_ws(m, a, b, nullptr, nullptr, nullptr);
    double glo = a[0];
    for (int i = 0; i < m; ++i) {
        const double r = (i ? std::fabs(b[i - 1]) : 0.0) + (i + 1 < m ? std::fabs(b[i]) : 0.0);
        glo = std::min(glo, a[i] - r);
    }
    const double scale = std::max(1.0, std::fabs(prev));
    const double hi0 = prev + 1e-13 * scale;
    if (sturm_count(m, a, b, hi0) < 1) return tridiag_lowest_ws(m, a, b, nullptr, nullptr, nullptr);
    double w = std::max(4.0 * step, 1e-12 * scale);
    double lo = prev - w;
    while (lo > glo && sturm_count(m, a, b, lo) >= 1) {
        w *= 16.0;
        lo = prev - w;
    }
    if (lo < glo) lo = glo;
    double hi = hi0;
    for (int it = 0; it < 200; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        if (sturm_count(m, a, b, mid) >= 1) hi = mid; else lo = mid;
    }
    return 0.5 * (lo + hi);
}

double tridiag_lowest(int m, const double* a, const double* b, double* vec) {
    if (!vec) return tridiag_lowest_ws(m, a, b, nullptr, nullptr, nullptr);
    std::vector<double> ws(8 * (size_t)m);
    std::vector<int> piv(m);
    return tridiag_lowest_ws(m, a, b, vec, ws.data(), piv.data());
}

}  // namespace blqh

// ---- kernel timing -----------------------------------------------------------------
namespace blqh {
Profile& profile() {
    static Profile p;
    return p;
}
}  // namespace blqh

// ---- public stage functions ------------------------------------------------------
void blqh_device_copy_impl(const void* src, void* dst, int64_t bytes) {
    BLQH_REQUIRE(bytes >= 0 && bytes % 16 == 0, BLQH_EINVAL, "copy size must be a multiple of 16 bytes");
    const int64_t n = bytes / 16;
    hipLaunchKernelGGL(k_copy, dim3(vec_grid(n)), dim3(kThreads), 0, blqh::stream(), (const double2*)src,
                       (double2*)dst, n);
    BLQH_HIP(hipGetLastError());
}

void blqh_spmv_device_impl(blqh_hamil* H, const void* x, void* y) {
    autotune(H);
    hipStream_t st = blqh::stream();
    if (!H->ord.on) {
        launch_spmv_plan(H, plan_of(H), (const double2*)x, (double2*)y, st);
        BLQH_HIP(hipGetLastError());
        return;
    }
    // x and y are in the reference order; the SpMV formats in the internal one
    DevBuf<double2> xi(H->n), yi(H->n);
    order_to_internal(H, (const double2*)x, xi.get(), st);
    launch_spmv_plan(H, plan_of(H), xi.get(), yi.get(), st);
    order_to_reference(H, yi.get(), (double2*)y, st);
    BLQH_HIP(hipGetLastError());
    BLQH_HIP(hipStreamSynchronize(st));  // xi / yi are freed on return
}

namespace {
__global__ void k_div(int64_t n, double2* __restrict__ x, double s) {
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        double2 v = x[i];
        v.x /= s;
        v.y /= s;
        x[i] = v;
    }
}

// Pinned host staging for start vectors, reused by every solve (grow-only,
// held until the process ends).
std::mutex g_stage_mu;
double2* g_stage = nullptr;
int64_t g_stage_n = 0;

// Start vector of pass 1 and pass 3 into device memory dst: the caller's, or
// the seeded mt19937 one of the reference (draw 2i -> imaginary, 2i+1 -> real
// part of entry i).  Drawn (or copied) and summed on the host in pinned
// memory, in the order of start_vector, then divided by the norm on the
// device: the same IEEE divisions as the host version, without a second
// host pass, fresh pages or a pageable copy per solve.
void device_start(int64_t n, uint32_t seed, const double* start, double2* dst, hipStream_t st,
                  double mix = 0.0) {
    std::lock_guard<std::mutex> lk(g_stage_mu);
    if (g_stage_n < n) {
        if (g_stage) (void)hipHostFree(g_stage);
        g_stage = nullptr;
        g_stage_n = 0;
        BLQH_HIP(hipHostMalloc((void**)&g_stage, (size_t)std::max<int64_t>(n, 1) * sizeof(double2),
                               hipHostMallocDefault));
        g_stage_n = n;
    }
    double s = 0;
    if (start && mix > 0) {
        std::vector<double2> r;
        blqh::start_vector(n, seed, r);  // unit vector
        const double2* z = (const double2*)start;
        double zs = 0;
        for (int64_t i = 0; i < n; ++i) zs += z[i].x * z[i].x + z[i].y * z[i].y;
        BLQH_REQUIRE(zs > 0 && std::isfinite(zs), BLQH_EINVAL, "start vector is zero or not finite");
        zs = 1.0 / std::sqrt(zs);
        for (int64_t i = 0; i < n; ++i) {
            g_stage[i] = make_double2(z[i].x * zs + mix * r[i].x, z[i].y * zs + mix * r[i].y);
            s += g_stage[i].x * g_stage[i].x + g_stage[i].y * g_stage[i].y;
        }
    } else if (start) {
        const double2* z = (const double2*)start;
        for (int64_t i = 0; i < n; ++i) {
            g_stage[i] = z[i];
            s += z[i].x * z[i].x + z[i].y * z[i].y;
        }
        BLQH_REQUIRE(s > 0 && std::isfinite(s), BLQH_EINVAL, "start vector is zero or not finite");
    } else {
        s = blqh::start_entries(n, seed, g_stage);
    }
    s = std::sqrt(s);
    BLQH_HIP(hipMemcpyAsync(dst, g_stage, n * sizeof(double2), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_div, dim3(vec_grid(n)), dim3(kThreads), 0, st, n, dst, s);
    BLQH_HIP(hipGetLastError());
    BLQH_HIP(hipStreamSynchronize(st));  // the staging buffer is free again
}
}  // namespace

namespace {

// Work areas of one solve.  Device: Krylov block (store: nvec = cap + 1
// vectors) or ring of three, u, psi, the replay start v0, the per-step
// scalars (3 cap + 8), coefficients (cap + 1), reductions (4), block
// partials of K1 and of K2 (npart each).  Host: arrays of cap + 1
// doubles and the tridiagonal work space (8 cap doubles, cap ints).
struct DevWork {
    double2* Wb;
    double2* u;
    double2* psi;
    double2* v0;
    double* sc;
    double* coef;
    double* red;
    double* partials;   // K1's
    double* partials2;  // K2's
    int64_t npart;
};

// K2 and K3 of one step after the K1 launched with grid `g1` wrote part1.
void launch_lz23(int64_t n, const double2* u, const double2* W, double2* Wn, const Scal& sc, int j, int ns,
                 double2* psi, const double* coef, int ldc, const double* part1, int g1, double* part2,
                 hipStream_t st) {
    const int gv = vec_grid(n);
    hipLaunchKernelGGL(k_lz2, dim3(gv), dim3(kThreads), 0, st, n, u, W, Wn, sc, j, ns, psi, coef, ldc, part1, g1,
                       part2);
    hipLaunchKernelGGL(k_lz3, dim3(1), dim3(kThreads), 0, st, sc, j, part2, gv);
}

// *out = |x|^2 and *out = Re<x|y>: block partials, then one wavefront sums them
void reduce_norm2(int64_t n, const double2* x, double* out, double* part, hipStream_t st) {
    const int gv = vec_grid(n);
    hipLaunchKernelGGL(k_norm2, dim3(gv), dim3(kThreads), 0, st, n, x, part);
    hipLaunchKernelGGL(k_psum, dim3(1), dim3(kThreads), 0, st, part, gv, out);
}
void reduce_redot(int64_t n, const double2* x, const double2* y, double* out, double* part, hipStream_t st) {
    const int gv = vec_grid(n);
    hipLaunchKernelGGL(k_redot, dim3(gv), dim3(kThreads), 0, st, n, x, y, part);
    hipLaunchKernelGGL(k_psum, dim3(1), dim3(kThreads), 0, st, part, gv, out);
}
struct HostWork {
    double *ha, *hb, *ritz, *svec, *c, *aux, *aux2, *tri;
    int* piv;
};

// Passes 1-3 and the Rayleigh quotient of blqh_lanczos on given work areas.
// start(dst) writes the normalised start vector into device memory dst.
// timed: bracket the fused SpMV launches with the profiling events.
template <class Start>
void lanczos_core(blqh_hamil* H, const Plan& plan, const blqh_lanczos_opts& o, bool store, const DevWork& D,
                  const HostWork& hw, Start&& start, bool timed, blqh_lanczos_result& res, double* psi_out,
                  double* alpha_out, double* beta_out, std::chrono::steady_clock::time_point t0) {
    hipStream_t st = stream();
    const int64_t n = H->n;
    const int cap = o.lambda;
    const int chk = std::max(1, o.check_every);
    const bool skip = o.skip_states != 0;
    const int gv = vec_grid(n);
    BLQH_REQUIRE(lz1_grid(H, plan) <= D.npart && spmv_grid(n, 4) <= D.npart && gv <= D.npart, BLQH_EINTERNAL,
                 "partial-sum buffer smaller than the grid");
    const int g1 = lz1_grid(H, plan);  // K1's partials, summed by K2
    Scal sc{D.sc, D.sc + cap, D.sc + 2 * cap + 2};
    const double one = 1.0;
    BLQH_HIP(hipMemcpyAsync(sc.beta, &one, 8, hipMemcpyHostToDevice, st));
    BLQH_HIP(hipMemcpyAsync(sc.invb, &one, 8, hipMemcpyHostToDevice, st));

    auto slot = [&](int j) -> double2* { return D.Wb + (size_t)(store ? j : j % 3) * n; };
    if (H->ord.on) {  // start vector in the reference order, the recurrence in the internal one
        start(D.u);
        order_to_internal(H, D.u, slot(0), st);
    } else {
        start(slot(0));
    }
    if (!store) BLQH_HIP(hipMemcpyAsync(D.v0, slot(0), n * sizeof(double2), hipMemcpyDeviceToDevice, st));

    Profile& prof = profile();
    const bool prof_on = timed && prof.on;
    int ntimed = 0;
    auto step = [&](int j, double2* psi_acc, const double* cf) {
        double2* W = slot(j);
        double2* Wp = j > 0 ? slot(j - 1) : slot(j);
        if (prof_on) {
            while ((int)prof.ev.size() < 2 * (ntimed + 1)) {
                hipEvent_t e;
                BLQH_HIP(hipEventCreate(&e));
                prof.ev.push_back(e);
            }
            BLQH_HIP(hipEventRecord(prof.ev[2 * ntimed], st));
        }
        launch_lz1(H, plan, W, Wp, D.u, sc, j, D.partials, st);
        if (prof_on) BLQH_HIP(hipEventRecord(prof.ev[2 * ntimed++ + 1], st));
        launch_lz23(n, D.u, W, slot(j + 1), sc, j, psi_acc ? 1 : 0, psi_acc, cf, cap, D.partials, g1, D.partials2,
                    st);
    };
    auto account = [&]() {
        if (!prof_on) return;
        for (int k = 0; k < ntimed; ++k) {
            float ms = 0;
            BLQH_HIP(hipEventElapsedTime(&ms, prof.ev[2 * k], prof.ev[2 * k + 1]));
            prof.ms += ms;
        }
        prof.launches += ntimed;
        prof.bytes += (double)ntimed * (20.0 * H->nnz + 56.0 * n + 8.0);
    };

    // pass 1
    double* ha = hw.ha;
    double* hb = hw.hb;
    int m = 0;  // Krylov vectors in use
    bool conv = false;
    double dE = INFINITY, resid = INFINITY;
    int done = 0;
    while (done < cap && !conv) {
        const int upto = std::min(cap, done + chk);
        for (int j = done; j < upto; ++j) step(j, nullptr, nullptr);
        BLQH_HIP(hipGetLastError());
        BLQH_HIP(hipMemcpyAsync(ha, sc.alpha, upto * 8, hipMemcpyDeviceToHost, st));
        BLQH_HIP(hipMemcpyAsync(hb, sc.beta, (upto + 1) * 8, hipMemcpyDeviceToHost, st));
        BLQH_HIP(hipStreamSynchronize(st));
        for (int j = done; j < upto; ++j) {
            // T_{j+1}: alpha[0..j], off-diagonal beta[1..j]
            hw.ritz[j] = j == 0 ? tridiag_lowest_ws(1, ha, hb + 1, nullptr, nullptr, nullptr)
                                : tridiag_lowest_next(j + 1, ha, hb + 1, hw.ritz[j - 1],
                                                      j > 1 ? std::fabs(hw.ritz[j - 1] - hw.ritz[j - 2]) : INFINITY);
            m = j + 1;
            if (j > 0) dE = std::fabs(hw.ritz[j] - hw.ritz[j - 1]);
            const double scale = std::max(std::fabs(ha[j]), 1.0);
            if (!(hb[j + 1] > 1e-13 * scale) || !std::isfinite(hb[j + 1])) {  // invariant subspace
                resid = 0;
                conv = true;
                break;
            }
            if (o.tol > 0 && j > 0 && dE < o.tol) {
                if (o.res_tol <= 0) { conv = true; break; }
                // residual of the Ritz pair: |H y - theta y| = beta_{j+1} |s_j|
                (void)tridiag_lowest_ws(j + 1, ha, hb + 1, hw.svec, hw.tri, hw.piv);
                resid = hb[j + 1] * std::fabs(hw.svec[j]);
                if (resid < o.res_tol) { conv = true; break; }
            }
        }
        done = upto;
    }
    BLQH_REQUIRE(std::isfinite(hw.ritz[m - 1]), BLQH_EINTERNAL, "Lanczos produced non-finite coefficients");

    // pass 2: lowest eigenpair of T_m
    double* c = hw.c;
    const double E0 = tridiag_lowest_ws(m, ha, hb + 1, c, hw.tri, hw.piv);
    if (alpha_out) { std::fill(alpha_out, alpha_out + cap, 0.0); std::copy(ha, ha + m, alpha_out); }
    if (beta_out) { std::fill(beta_out, beta_out + cap + 1, 0.0); std::copy(hb, hb + m + 1, beta_out); }
    res.E0 = E0;
    res.dE = dE;
    res.residual = std::isfinite(resid) ? resid : hb[m] * std::fabs(c[m - 1]);
    res.iterations = m;
    res.converged = conv ? 1 : 0;
    if (skip) {
        res.E_rq = NAN;
        res.passes = 1;
        BLQH_HIP(hipStreamSynchronize(st));
        account();
        res.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return;
    }
    // pass 3
    if (store) {
        double* invb = hw.aux;
        copy_sync(invb, sc.invb, m * 8, hipMemcpyDeviceToHost);
        double* w = hw.aux2;
        for (int l = 0; l < m; ++l) w[l] = c[l] * invb[l];
        BLQH_HIP(hipMemcpyAsync(D.coef, w, m * 8, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_combine, dim3(gv), dim3(kThreads), 0, st, n, m, D.Wb, n, D.coef, D.psi);
        res.passes = 2;
    } else {
        // replay with the same kernels: alpha/beta are regenerated bit for bit
        BLQH_HIP(hipMemcpyAsync(D.coef, c, m * 8, hipMemcpyHostToDevice, st));
        BLQH_HIP(hipMemsetAsync(D.psi, 0, n * sizeof(double2), st));
        BLQH_HIP(hipMemcpyAsync(slot(0), D.v0, n * sizeof(double2), hipMemcpyDeviceToDevice, st));
        for (int j = 0; j + 1 < m; ++j) step(j, D.psi, D.coef);
        // last vector's weight needs invb[m-1], which the replay just recomputed
        double* hb2 = hw.aux;
        double* ha2 = hw.aux2;
        BLQH_HIP(hipMemcpyAsync(hb2, sc.invb, m * 8, hipMemcpyDeviceToHost, st));
        if (m > 1) BLQH_HIP(hipMemcpyAsync(ha2, sc.alpha, (m - 1) * 8, hipMemcpyDeviceToHost, st));
        BLQH_HIP(hipStreamSynchronize(st));
        for (int j = 0; j + 1 < m; ++j)
            BLQH_REQUIRE(ha2[j] == ha[j], BLQH_EINTERNAL, "pass-3 replay diverged from pass 1");
        hipLaunchKernelGGL(k_accum, dim3(gv), dim3(kThreads), 0, st, n, slot(m - 1), c[m - 1] * hb2[m - 1], D.psi);
        res.passes = 3;
    }
    // normalise and Rayleigh quotient
    reduce_norm2(n, D.psi, D.red, D.partials, st);
    hipLaunchKernelGGL(k_scale, dim3(gv), dim3(kThreads), 0, st, n, D.psi, D.red);
    launch_spmv_plan(H, plan, D.psi, D.u, st);
    reduce_redot(n, D.psi, D.u, D.red + 1, D.partials, st);
    BLQH_HIP(hipGetLastError());
    double erq = 0;
    BLQH_HIP(hipMemcpyAsync(&erq, D.red + 1, 8, hipMemcpyDeviceToHost, st));
    if (psi_out) {
        const double2* out = D.psi;
        if (H->ord.on) {
            order_to_reference(H, D.psi, D.u, st);
            out = D.u;
        }
        BLQH_HIP(hipMemcpyAsync(psi_out, out, n * sizeof(double2), hipMemcpyDeviceToHost, st));
    }
    BLQH_HIP(hipStreamSynchronize(st));
    account();
    res.E_rq = erq;
    res.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

void blqh_lanczos_impl(blqh_hamil* H, const blqh_lanczos_opts& o, blqh_lanczos_result& res, double* psi_out,
                       double* alpha_out, double* beta_out) {
    using namespace blqh;
    auto t0 = std::chrono::steady_clock::now();
    hipStream_t st = stream();
    const int64_t n = H->n;
    const int cap = o.lambda;
    BLQH_REQUIRE(cap >= 1 && cap <= 100000, BLQH_EINVAL, "lambda out of range");
    BLQH_REQUIRE(o.start_mix >= 0 && std::isfinite(o.start_mix), BLQH_EINVAL, "start_mix must be finite and >= 0");
    autotune(H);
    const Plan plan = plan_of(H);

    // Krylov storage: keep all vectors when they fit comfortably
    size_t free_b = 0, total_b = 0;
    BLQH_HIP(hipMemGetInfo(&free_b, &total_b));
    free_b += dev_cached_bytes();  // cached blocks are released if an allocation needs them
    const size_t need = (size_t)(cap + 1) * n * sizeof(double2);
    const bool skip = o.skip_states != 0;
    const bool store = !skip && (o.store_basis > 0 || (o.store_basis < 0 && need < free_b / 2));
    const int nvec = store ? cap + 1 : 3;
    DevBuf<double2> Wb((size_t)nvec * n), u(n), psi(n), v0(store ? 0 : n);
    DevBuf<double> sc_buf(3 * (size_t)cap + 8), coef(cap + 1), red(4);
    const int gmax = std::max(std::max(lz1_grid(H, plan), spmv_grid(n, 4)), vec_grid(n));
    DevBuf<double> partials(2 * (size_t)gmax);
    std::vector<double> hbuf(7 * ((size_t)cap + 1) + 8 * (size_t)cap);
    std::vector<int> piv(cap);
    double* hp = hbuf.data();
    const size_t L = (size_t)cap + 1;
    const HostWork hw{hp, hp + L, hp + 2 * L, hp + 3 * L, hp + 4 * L, hp + 5 * L, hp + 6 * L, hp + 7 * L, piv.data()};
    const DevWork D{Wb.get(), u.get(), psi.get(), v0.get(), sc_buf.get(), coef.get(), red.get(), partials.get(),
                    partials.get() + gmax, gmax};
    lanczos_core(H, plan, o, store, D, hw,
                 [&](double2* dst) { device_start(n, o.seed, o.start, dst, st, o.start_mix); }, true, res, psi_out,
                 alpha_out, beta_out, t0);
}

// ---- the allocation-free solve (blqh_lanczos_noalloc) ----------------------------
//
// Every device array of the solve lives in `scratch`, one static array per
// GPU of the library, sized for kNoAllocMaxRows rows and kNoAllocMaxLambda
// steps.  Host code never names the array itself: blqh::noalloc_setup (run by
// the library's device initialisation, with the device current) asks the
// runtime for this device's address and every use goes through that pointer.
// Calls are serialised by a mutex (they share the array); each call
// initialises everything it reads (scalars, psi), so
// nothing carries over from the previous call's contents.
namespace {
constexpr int64_t kNoAllocMaxRows = int64_t(1) << 17;
constexpr int kNoAllocMaxLambda = 1024;
constexpr int64_t kNoAllocPartials = 16384;
constexpr size_t kNoAllocVecBytes = (size_t)kNoAllocMaxRows * sizeof(double2);
// ring of three Krylov vectors, u, psi, v0
constexpr size_t kNoAllocScalOff = 6 * kNoAllocVecBytes;
constexpr size_t kNoAllocScalDoubles = (3 * (size_t)kNoAllocMaxLambda + 8) + ((size_t)kNoAllocMaxLambda + 1) + 4 +
                                       (size_t)kNoAllocPartials;
constexpr size_t kNoAllocBytes = kNoAllocScalOff + kNoAllocScalDoubles * sizeof(double) + 256;
constexpr int kMaxDevices = 64;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {
std::mutex g_noalloc_mu;
void* g_scratch_ptr[kMaxDevices] = {};
// host work arrays of the allocation-free solve (used under g_noalloc_mu)
double g_noalloc_host[7 * (kNoAllocMaxLambda + 1) + 8 * kNoAllocMaxLambda];
int g_noalloc_piv[kNoAllocMaxLambda];

// The seeded start vector drawn on the device (mt_start_block, one workgroup).
__global__ void __launch_bounds__(1024) k_mt_start(int64_t n, uint32_t seed, double2* __restrict__ v,
                                                   double* __restrict__ nrm2) {
    mt_start_block(n, seed, v, nrm2);
}

// The SpMV plan the Hamiltonian's current layout supports without any
// conversion (the tuned one when its layout is in place, else the CSR
// kernel, whose arrays always exist).
Plan noalloc_plan(const blqh_hamil* H) {
    if (H->spmv_kind >= 0) {
        const Plan p = plan_of(H);
        const Panels& P = H->pan;
        bool ok = false;
        switch (p.kind) {
            case 0: ok = true; break;
            case 1:
            case 2: ok = P.built && !P.jagged; break;
            case 3: ok = P.built && P.jagged && P.greedy == greedy_step_order() && P.rows == 1; break;
            case 4: ok = H->til.built; break;
            case 5: ok = P.built && P.jagged && P.rows == p.lpr; break;
            default: break;
        }
        if (ok) return p;
    }
    const double mean = H->n ? (double)H->nnz / (double)H->n : 0.0;
    int lpr = 4;
    while (lpr < 64 && 4.0 * lpr < mean) lpr *= 2;
    return Plan{0, lpr, kThreads, 0, 0};
}
}  // namespace

namespace blqh {
// This device's address of `scratch` (the library's device initialisation
// calls it once per device, with that device current).
void noalloc_setup(int device) {
    BLQH_REQUIRE(device >= 0 && device < kMaxDevices, BLQH_EINVAL, "device index out of range");
    void* p = nullptr;
    BLQH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    BLQH_REQUIRE(p != nullptr, BLQH_EDEVICE, "no device address for the solver's static scratch array");
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    g_scratch_ptr[device] = p;
}
}  // namespace blqh

void blqh_lanczos_noalloc_impl(blqh_hamil* H, const blqh_lanczos_opts& o, blqh_lanczos_result& res, double* psi_out,
                               double* alpha_out, double* beta_out) {
    using namespace blqh;
    auto t0 = std::chrono::steady_clock::now();
    hipStream_t st = stream();
    const int64_t n = H->n;
    const int cap = o.lambda;
    BLQH_REQUIRE(cap >= 1, BLQH_EINVAL, "lambda out of range");
    BLQH_REQUIRE(cap <= kNoAllocMaxLambda, BLQH_ERANGE, "blqh_lanczos_noalloc: lambda above 1024");
    BLQH_REQUIRE(n >= 1 && n <= kNoAllocMaxRows, BLQH_ERANGE, "blqh_lanczos_noalloc: more than 131072 rows");
    BLQH_REQUIRE(o.start_mix == 0.0, BLQH_EINVAL, "blqh_lanczos_noalloc: start_mix is not supported");
    const Plan plan = noalloc_plan(H);
    const int dev = runtime().device;  // initialised above (stream()), outside the lock: set-up takes it too
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    unsigned char* base = static_cast<unsigned char*>(g_scratch_ptr[dev]);
    BLQH_REQUIRE(base != nullptr, BLQH_EINTERNAL, "solver scratch not set up for this device");
    double2* vec = reinterpret_cast<double2*>(base);
    double* sd = reinterpret_cast<double*>(base + kNoAllocScalOff);
    // scalars, coefficients, reductions and partials packed for this lambda
    double* sc = sd;
    double* coef = sc + 3 * (size_t)cap + 8;
    double* red = coef + cap + 1;
    double* partials = red + 4;
    const DevWork D{vec, vec + 3 * n, vec + 4 * n, vec + 5 * n, sc, coef, red, partials,
                    partials + kNoAllocPartials / 2, kNoAllocPartials / 2};
    double* hp = g_noalloc_host;
    const size_t L = (size_t)cap + 1;
    const HostWork hw{hp, hp + L, hp + 2 * L, hp + 3 * L, hp + 4 * L, hp + 5 * L, hp + 6 * L, hp + 7 * L,
                      g_noalloc_piv};
    auto start = [&](double2* dst) {
        double s = 0;
        if (o.start) {
            BLQH_HIP(hipMemcpyAsync(dst, o.start, n * sizeof(double2), hipMemcpyHostToDevice, st));
            const double2* z = reinterpret_cast<const double2*>(o.start);
            for (int64_t i = 0; i < n; ++i) s += z[i].x * z[i].x + z[i].y * z[i].y;
            BLQH_REQUIRE(s > 0 && std::isfinite(s), BLQH_EINVAL, "start vector is zero or not finite");
        } else {
            hipLaunchKernelGGL(k_mt_start, dim3(1), dim3(1024), 0, st, n, o.seed, dst, red + 2);
            BLQH_HIP(hipGetLastError());
            copy_sync(&s, red + 2, sizeof(double), hipMemcpyDeviceToHost);
        }
        hipLaunchKernelGGL(k_div, dim3(vec_grid(n)), dim3(kThreads), 0, st, n, dst, std::sqrt(s));
        BLQH_HIP(hipGetLastError());
    };
    lanczos_core(H, plan, o, false, D, hw, start, false, res, psi_out, alpha_out, beta_out, t0);
}

void blqh_lanczos_reconstruct_impl(blqh_hamil* H, int m, uint32_t seed, const double* start, const double* alpha,
                                   const double* beta, int ns, const double* coefs, double* psi_out, double* erq_out,
                                   double* dev_out) {
    using namespace blqh;
    hipStream_t st = stream();
    const int64_t n = H->n;
    BLQH_REQUIRE(m >= 1 && m <= 100000, BLQH_EINVAL, "lambda out of range");
    BLQH_REQUIRE(ns >= 1 && ns <= 16, BLQH_EINVAL, "between 1 and 16 states per replay");
    BLQH_REQUIRE(coefs != nullptr, BLQH_EINVAL, "coefficients missing");
    for (int s = 0; s < ns; ++s) {
        bool any = false;
        for (int l = 0; l < m; ++l) {
            const double c = coefs[(size_t)s * m + l];
            BLQH_REQUIRE(std::isfinite(c), BLQH_EINVAL, "coefficients not finite");
            any |= c != 0.0;
        }
        BLQH_REQUIRE(any, BLQH_EINVAL, "a coefficient vector is zero");
    }
    autotune(H);
    const Plan plan = plan_of(H);
    DevBuf<double2> Wb(3 * (size_t)n), u(n), psi((size_t)ns * n);
    DevBuf<double> sc_buf(3 * (size_t)m + 8), coef((size_t)ns * m), red(2 * ns + 2);
    const int gmax = std::max(std::max(lz1_grid(H, plan), spmv_grid(n, 4)), vec_grid(n));
    DevBuf<double> partials(2 * (size_t)gmax);
    Scal sc{sc_buf.get(), sc_buf.get() + m, sc_buf.get() + 2 * m + 2};
    const double one = 1.0;
    BLQH_HIP(hipMemcpyAsync(sc.beta, &one, 8, hipMemcpyHostToDevice, st));
    BLQH_HIP(hipMemcpyAsync(sc.invb, &one, 8, hipMemcpyHostToDevice, st));

    auto slot = [&](int j) -> double2* { return Wb.get() + (size_t)(j % 3) * n; };
    if (H->ord.on) {
        device_start(n, seed, start, u.get(), st);
        order_to_internal(H, u.get(), slot(0), st);
    } else {
        device_start(n, seed, start, slot(0), st);
    }
    BLQH_HIP(hipMemcpyAsync(coef.get(), coefs, (size_t)ns * m * 8, hipMemcpyHostToDevice, st));
    BLQH_HIP(hipMemsetAsync(psi.get(), 0, (size_t)ns * n * sizeof(double2), st));
    const int gv = vec_grid(n);
    for (int j = 0; j + 1 < m; ++j) {
        double2* W = slot(j);
        double2* Wp = j > 0 ? slot(j - 1) : slot(j);
        launch_lz1(H, plan, W, Wp, u.get(), sc, j, partials.get(), st);
        launch_lz23(n, u.get(), W, slot(j + 1), sc, j, ns, psi.get(), coef.get(), m, partials.get(), lz1_grid(H, plan),
                    partials.get() + gmax, st);
    }
    BLQH_HIP(hipGetLastError());
    std::vector<double> ra(m), rb(m + 1);
    BLQH_HIP(hipMemcpyAsync(ra.data(), sc.alpha, m * 8, hipMemcpyDeviceToHost, st));
    BLQH_HIP(hipMemcpyAsync(rb.data(), sc.invb, m * 8, hipMemcpyDeviceToHost, st));
    std::vector<double> rbeta(m);
    BLQH_HIP(hipMemcpyAsync(rbeta.data(), sc.beta, m * 8, hipMemcpyDeviceToHost, st));
    BLQH_HIP(hipStreamSynchronize(st));
    for (int s = 0; s < ns; ++s)  // the last vector v_{m-1} = invb[m-1] W_{m-1}
        hipLaunchKernelGGL(k_accum, dim3(gv), dim3(kThreads), 0, st, n, slot(m - 1),
                           coefs[(size_t)s * m + m - 1] * rb[m - 1], psi.get() + (size_t)s * n);
    double dev = 0.0;
    for (int j = 0; j + 1 < m; ++j) {
        if (alpha) dev = std::max(dev, std::fabs(ra[j] - alpha[j]) / std::max(1.0, std::fabs(alpha[j])));
        if (beta) dev = std::max(dev, std::fabs(rbeta[j + 1] - beta[j + 1]) / std::max(1.0, std::fabs(beta[j + 1])));
    }
    BLQH_REQUIRE(std::isfinite(dev), BLQH_EINTERNAL, "replay produced non-finite coefficients");
    for (int s = 0; s < ns; ++s) {
        double2* p = psi.get() + (size_t)s * n;
        reduce_norm2(n, p, red.get() + 2 * s, partials.get(), st);
        hipLaunchKernelGGL(k_scale, dim3(gv), dim3(kThreads), 0, st, n, p, red.get() + 2 * s);
        launch_spmv_plan(H, plan, p, u.get(), st);
        reduce_redot(n, p, u.get(), red.get() + 2 * s + 1, partials.get(), st);
    }
    BLQH_HIP(hipGetLastError());
    std::vector<double> hred(2 * ns);
    BLQH_HIP(hipMemcpyAsync(hred.data(), red.get(), 2 * ns * 8, hipMemcpyDeviceToHost, st));
    if (psi_out && H->ord.on) {
        for (int s = 0; s < ns; ++s) {  // back to the reference order, through u
            order_to_reference(H, psi.get() + (size_t)s * n, u.get(), st);
            BLQH_HIP(hipMemcpyAsync(psi_out + (size_t)s * 2 * n, u.get(), n * sizeof(double2), hipMemcpyDeviceToHost,
                                    st));
        }
    } else if (psi_out) {
        BLQH_HIP(hipMemcpyAsync(psi_out, psi.get(), (size_t)ns * n * sizeof(double2), hipMemcpyDeviceToHost, st));
    }
    BLQH_HIP(hipStreamSynchronize(st));
    for (int s = 0; s < ns; ++s) {
        BLQH_REQUIRE(hred[2 * s] > 0, BLQH_EINVAL, "a reconstructed state is zero (coefficients all zero?)");
        if (erq_out) erq_out[s] = hred[2 * s + 1];
    }
    if (dev_out) *dev_out = dev;
}
