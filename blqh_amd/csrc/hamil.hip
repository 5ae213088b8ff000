// hamil.hip -- Hamiltonian build into CSR on the GPU.
//
// Reference: lhamil::init_Coulomb_matrix + peer_set_hamil + set_hamil,
// BLQH_without_spin/lanczos_hamiltonian.cpp:37-385 (spinless) and
// BLQH_with_spin/lanczos_hamiltonian.cpp:37-954 (spinful).
//
// Row i of H (one wavefront per row) sums the reference's terms: for every
//   translation L_ql of id[i] (ql < D_l), every occupied pair (n < m) and
//   every shift t in (-nphi, nphi): n -> nt = n + t, m -> mt = m - t (both
//   wrapped inside their species block); when both targets are empty the
//   target word R0 is looked up through its combinatorial rank in the sector
//   table, which gives the representative row j, the shift qr with
//   inv_translate(R0, qr) = id[j] and that translation's sign, and the
//   element FT[ql][qr] * V[alpha][|dm|][|t|] * sign / sqrt(D_l D_r) goes into
//   an LDS hash keyed by j.  The kernel visits them per open hop of id[i]:
//   the translations map those hops one to one onto the hops of every L_ql,
//   with targets in the same orbit, so one lookup serves the D_l images,
//   whose sum is inserted once.  Tunnelling terms are one-body hops between
//   the layers; the diagonal (background, charging, bias, Zeeman) comes last.
// Entries with |h_ij| > 1e-10 are kept, columns ascending.
//
// Default: one staged pass.  When a fixed-size staging slot per row fits in a
// quarter of free device memory, each row is hashed, sorted and written to
// its slot with its length; the lengths are scanned into rowptr and a copy
// kernel compacts the slots into the CSR.  Otherwise the build runs twice over
// the rows: a count pass (row lengths), an exclusive scan into rowptr, and a
// fill pass that writes columns and values.  Each row is owned by exactly one
// wavefront, so its LDS accumulation order is fixed and both schemes produce
// the same entries bit for bit; the fill pass checks this.  Rows whose
// distinct columns overflow the small hash are redone by a second kernel
// instance with a large hash.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <complex>
#include <mutex>
#include <thread>

#include "engine.hpp"
#include "rowterms.hpp"

namespace blqh {

// Coulomb tables, background energy and translation phases.
// BLQH_without_spin/lanczos_hamiltonian.cpp:37-85 (theta_x enters as +2 theta_x);
// BLQH_with_spin/lanczos_hamiltonian.cpp:37-84 (as -theta_x).
static double vq(const blqh_params& p, int alpha, int qx, int qy) {
    const double q = std::sqrt(qx * qx / (p.lx * p.lx) + qy * qy / (p.ly * p.ly)) * 2.0 * M_PI;
    if (alpha == 1)
        return 2.0 * M_PI / q * std::exp(-q * q / 2.0 - q * p.d) * std::pow(1.0 - std::exp(-q * q / 2.0), p.nLL * 2);
    return 2.0 * M_PI / q * std::exp(-q * q / 2.0) * std::pow(1.0 - std::exp(-q * q / 2.0), p.nLL * 2);
}

// The Coulomb tables depend on the geometry and d only, not on the sector,
// so a sweep over all (J,K) sectors at one parameter point computes them once.
// The interlayer lattice sum runs over ~20 nphi/(d+0.01) terms per entry
// (24,000 at d = 0, nphi = 12); entries are spread over threads, each entry
// summed in the reference's order.
namespace {
struct CoulombKey {
    int nphi, nspecies, nLL;
    double lx, ly, d, theta_x;
    bool operator==(const CoulombKey& o) const {
        return nphi == o.nphi && nspecies == o.nspecies && nLL == o.nLL && lx == o.lx && ly == o.ly && d == o.d &&
               theta_x == o.theta_x;
    }
};
struct CoulombEntry {
    CoulombKey key;
    std::vector<double> C;
    double Ec;
};
std::mutex g_coulomb_mu;
std::vector<CoulombEntry> g_coulomb;  // most recent last, at most 8

// Monolayer (Monolayer/lanczos_hamiltonian.cpp:53-76): V(sx, ky) sums kx over
// [-(nphi/2), nphi/2] (C integer division) and the element is 2 V(|dm|, |t|)
// for t in [-(nphi/2), nphi/2].  The row kernel runs t over (-nphi, nphi), so
// every target of shift t is met twice (at t and at t -/+ nphi) except t = 0;
// the table is folded to give the same sums: T(sx, 0) = 2 V(sx, 0), for even
// nphi T(sx, nphi/2) = 2 V(sx, nphi/2) (the tree visits both t = +-nphi/2),
// else T(sx, a) = V(sx, min(a, nphi - a)).  Background
// E_cl = (-2 + sum_{(i,j) in [0,nphi)^2 \ 0} sqrt(pi/z) erfc(sqrt(z))) / sqrt(lx ly),
// z = pi (i^2 lx/ly + j^2 ly/lx): the closed form of the tree's integral
// int_1^inf t^-1/2 e^{-zt} dt.  The interlayer half of the table stays zero.
void monolayer_tables(const blqh_sector& s, const blqh_params& p, std::vector<double>& C, double& Ec) {
    const int nphi = s.nphi;
    const double lx = p.lx, ly = p.ly;
    auto v = [&](int kx, int ky) {
        const double q = std::sqrt(kx * kx / (lx * lx) + ky * ky / (ly * ly)) * 2.0 * M_PI;
        return 2.0 * M_PI / (q + 1e-30) * std::exp(-q * q / 2.0) * std::pow(1.0 - std::exp(-q * q / 2.0), p.nLL * 2);
    };
    std::vector<double> V(nphi * nphi, 0.0);
    for (int sx = 0; sx < nphi; ++sx)
        for (int ky = 0; ky <= nphi / 2; ++ky) {
            double x = 0;
            for (int kx = -nphi / 2; kx <= nphi / 2; ++kx)
                if (!(ky == 0 && kx == 0)) x += v(kx, ky) * std::cos(2.0 * M_PI * sx * kx / nphi) / (2.0 * lx * ly);
            V[sx * nphi + ky] = x;
        }
    C.assign(2 * nphi * nphi, 0.0);
    for (int sx = 0; sx < nphi; ++sx)
        for (int a = 0; a < nphi; ++a) {
            const bool both = a == 0 || 2 * a == nphi;
            const int ky = a <= nphi - a ? a : nphi - a;
            C[sx * nphi + a] = both ? 2.0 * V[sx * nphi + ky] : V[sx * nphi + ky];
        }
    Ec = -2.0;
    for (int i = 0; i < nphi; ++i)
        for (int j = 0; j < nphi; ++j)
            if (!(i == 0 && j == 0)) {
                const double z = (i * i * lx / ly + j * j * ly / lx) * M_PI;
                Ec += std::sqrt(M_PI / z) * std::erfc(std::sqrt(z));
            }
    Ec /= std::sqrt(lx * ly);
}

void coulomb_tables(const blqh_sector& s, const blqh_params& p, std::vector<double>& C, double& Ec) {
    const CoulombKey key{s.nphi, s.nspecies, p.nLL, p.lx, p.ly, p.d, p.theta_x};
    {
        std::lock_guard<std::mutex> lk(g_coulomb_mu);
        for (const auto& e : g_coulomb)
            if (e.key == key) {
                C = e.C;
                Ec = e.Ec;
                return;
            }
    }
    const int nphi = s.nphi;
    const double lx = p.lx, ly = p.ly;
    if (s.nspecies == 1) {
        monolayer_tables(s, p, C, Ec);
        std::lock_guard<std::mutex> lk(g_coulomb_mu);
        if (g_coulomb.size() >= 8) g_coulomb.erase(g_coulomb.begin());
        g_coulomb.push_back({key, C, Ec});
        return;
    }
    Ec = -2.0;
    for (int i = 0; i < nphi; ++i)
        for (int j = 0; j < nphi; ++j)
            if (!(i == 0 && j == 0)) Ec += std::erfc(std::sqrt(M_PI * (i * i * lx / ly + j * j * ly / lx)));
    Ec /= std::sqrt(lx * ly);
    C.assign(2 * nphi * nphi, 0.0);
    auto entry = [&](int e) {
        const int sx = e / nphi, qy = e % nphi;
        double V = 0;
        for (int qx = -nphi; qx <= nphi; ++qx)
            if (!(qx == 0 && qy == 0)) V += 2.0 * vq(p, 0, qx, qy) * std::cos(2.0 * M_PI * sx * qx / nphi) / (2.0 * lx * ly);
        C[sx * nphi + qy] = V;
        V = 0;
        const double arg = s.nspecies == 2 ? 2.0 * M_PI * sx + 2 * p.theta_x : 2.0 * M_PI * sx - p.theta_x;
        const int q0 = (int)(-10 * nphi / (p.d + 0.01));
        for (int qx = q0; qx < 10 * nphi / (p.d + 0.01); ++qx)
            if (!(qx == 0 && qy == 0)) V += 2.0 * vq(p, 1, qx, qy) * std::cos(arg * qx / nphi) / (2.0 * lx * ly);
        C[nphi * nphi + sx * nphi + qy] = V;
    };
    const int ne = nphi * nphi;
    const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
            for (int e = t; e < ne; e += nt) entry(e);
        });
    for (auto& t : th) t.join();
    std::lock_guard<std::mutex> lk(g_coulomb_mu);
    if (g_coulomb.size() >= 8) g_coulomb.erase(g_coulomb.begin());
    g_coulomb.push_back({key, C, Ec});
}
}  // namespace

HostTables host_tables(const blqh_sector& s, const blqh_params& p) {
    HostTables H;
    const int nphi = s.nphi;
    if (s.nspecies == 1)
        BLQH_REQUIRE(p.d == 0 && p.delta_sas == 0 && p.delta_v == 0 && p.delta_z == 0 && p.theta_x == 0 &&
                         p.theta_y == 0 && p.theta_b == 0,
                     BLQH_EINVAL, "a monolayer takes no layer, Zeeman or twist parameters (d, Delta_*, theta_* = 0)");
    coulomb_tables(s, p, H.C, H.Ec);
    // the reference's FT[a][b] depends on a - b only (same argument, same bits);
    // the monolayer's phase is exp(2 pi i K dq / gcd(nel, nphi)) with |dq| < gcd
    H.FT.resize(2 * nphi - 1);
    const int ord = make_geo(s).order();
    for (int dq = -nphi + 1; dq < nphi; ++dq) {
        const double ang = s.nspecies == 1 ? 2.0 * M_PI * s.K * dq / ord : 2.0 * M_PI * dq * s.K / nphi;
        H.FT[dq + nphi - 1] = make_double2(std::cos(ang), std::sin(ang));
    }
    H.tw.resize(2 * nphi - 1);
    for (int t = -nphi + 1; t < nphi; ++t) {
        const double ang = s.nspecies == 2 ? 2 * p.theta_y * t / nphi : p.theta_y * t / nphi;
        H.tw[t + nphi - 1] = make_double2(std::cos(ang), std::sin(ang));
    }
    if (p.delta_sas > 0) {
        if (s.nspecies == 2) {
            // BLQH_without_spin/lanczos_hamiltonian.cpp:295-322
            for (int n = 0; n < nphi; ++n) H.hops.push_back({n, n + nphi, n, n + nphi, -0.5 * p.delta_sas, 1, 0, 0});
        } else {
            // BLQH_with_spin/lanczos_hamiltonian.cpp:718-898: same-spin hops appear in two
            // loops with -Delta_SAS/4 each (and theta_B phases), spin-flip hops once with -Delta_SAS/2.
            const double tb = p.theta_b, d = p.d;
            for (int n = 0; n < nphi; ++n)
                H.hops.push_back({n, n + 2 * nphi, n, n + 2 * nphi, -0.25 * p.delta_sas, std::cos(tb * n * d), std::sin(tb * n * d), 1});
            for (int n = 2 * nphi; n < 3 * nphi; ++n) {
                const int nt = n - 2 * nphi;
                H.hops.push_back({n, nt, nt, n, -0.25 * p.delta_sas, std::cos(tb * nt * d), -std::sin(tb * nt * d), 1});
            }
            for (int n = nphi; n < 2 * nphi; ++n)
                H.hops.push_back({n, n + 2 * nphi, n, n + 2 * nphi, -0.25 * p.delta_sas, std::cos(tb * (n % nphi) * d),
                                  std::sin(tb * (n % nphi) * d), 1});
            for (int n = 3 * nphi; n < 4 * nphi; ++n)
                H.hops.push_back({n, n - 2 * nphi, n - 2 * nphi, n, -0.25 * p.delta_sas, std::cos(tb * (n % nphi) * d),
                                  -std::sin(tb * (n % nphi) * d), 1});
            for (int n = 0; n < nphi; ++n) H.hops.push_back({n, n + 3 * nphi, n, n + 3 * nphi, -0.5 * p.delta_sas, 1, 0, 0});
            for (int n = nphi; n < 2 * nphi; ++n) H.hops.push_back({n, n + nphi, n, n + nphi, -0.5 * p.delta_sas, 1, 0, 0});
        }
    }
    return H;
}

}  // namespace blqh

namespace {

using namespace blqh;

struct BuildArgs {
    RowCtx ctx;                  // C / FT / tw point to global copies; the kernel restages them in LDS
    int64_t nrows;               // rows to process (size of row_list or n)
    const int32_t* row_list;     // optional row list (overflow pass)
    const uint64_t* id;
    const HopTerm* hops;
    int nhop;
    double Ecn, d, dv, dz;
    int32_t* nnz_out;         // count pass
    const int64_t* rowptr;    // fill pass
    int32_t* col;
    double2* val;
    int32_t* row_counter;
    int32_t* ovf_count;
    int32_t* ovf_rows;
    int32_t* err;             // fill/count mismatch counter
    int32_t* row_max;         // largest row (entries kept) seen by the count / single pass
    int32_t* stage_col;       // single pass: row i's sorted entries at i * CAP
    double2* stage_val;
    uint8_t* stage_ok;        // single pass: 1 when row i was staged (0: overflow row)
    int terms;                // bit 0: pair terms and diagonal; tunnelling hops come with the hop list
};

template <int CAP>
__device__ __forceinline__ bool hash_add(int* keys, double* vre, double* vim, int j, double re, double im) {
    constexpr int LOG = __builtin_ctz(CAP);
    unsigned h = ((unsigned)j * 2654435761u) >> (32 - LOG);
    for (int probe = 0; probe < CAP; ++probe) {
        int k = __atomic_load_n(&keys[h], __ATOMIC_RELAXED);
        if (k == -1) {
            const int old = atomicCAS(&keys[h], -1, j);
            k = old == -1 ? j : old;
        }
        if (k == j) {
            atomicAdd(&vre[h], re);
            atomicAdd(&vim[h], im);
            return true;
        }
        h = (h + 1) & (CAP - 1);
    }
    return false;
}

// MODE 0: count, MODE 1: fill
template <int CAP, int MODE>
__global__ void __launch_bounds__(64) k_build(BuildArgs A) {
    extern __shared__ double smem[];
    __shared__ int keys[CAP];
    __shared__ double vre[CAP], vim[CAP];
    __shared__ unsigned long long sbuf[MODE >= 1 && CAP > 512 ? CAP : 1];
    __shared__ int occ[64];
    __shared__ int sQ[128];  // open pair hops awaiting evaluation
    __shared__ uint64_t sLw[32];  // translate(L0, ql)
    __shared__ int sLs[32];       // and its sign
    __shared__ int s_row, s_ovf;

    const int lane = threadIdx.x;
    const Geo g = A.ctx.g;
    const int nphi = g.nphi, nbits = g.S * nphi, nel = g.nel;
    const int T = 2 * nphi - 1;
    // stage the tables in LDS
    double* sC = smem;                                   // 2 nphi^2
    double2* sFT = (double2*)(sC + 2 * nphi * nphi);     // T
    double2* sTW = sFT + T;                              // T
    unsigned short* sPair = (unsigned short*)(sTW + T);  // npairs
    const int npairs = nel * (nel - 1) / 2;
    for (int x = lane; x < 2 * nphi * nphi; x += 64) sC[x] = A.ctx.C[x];
    for (int x = lane; x < T; x += 64) sFT[x] = A.ctx.FT[x];
    for (int x = lane; x < T; x += 64) sTW[x] = A.ctx.tw[x];
    for (int x = lane; x < npairs; x += 64) sPair[x] = pair_decode(nel, x);
    RowCtx ctx = A.ctx;
    ctx.C = sC;
    ctx.FT = sFT;
    ctx.tw = sTW;
    __syncthreads();

    for (;;) {
        if (lane == 0) s_row = atomicAdd(A.row_counter, 1);
        __syncthreads();
        const int64_t r = s_row;
        if (r >= A.nrows) break;
        const int64_t i = A.row_list ? A.row_list[r] : r;
        for (int x = lane; x < CAP; x += 64) { keys[x] = -1; vre[x] = 0.0; vim[x] = 0.0; }
        if (lane == 0) s_ovf = 0;
        __syncthreads();
        bool ovf = false;
        const uint64_t L0 = A.id[i];
        const int Dl = ctx.K < 0 ? 1 : ctx.period[i];
        if (lane < Dl) {  // Dl <= nphi < 32
            int sg;
            sLw[lane] = translate(g, L0, lane, sg);
            sLs[lane] = sg;
        }
        if (lane < nbits && ((L0 >> lane) & 1ULL)) occ[popc(L0 & lowbits(lane))] = lane;
        __syncthreads();
        {
            // Pair hops of the row word L0: candidate cnd = pr * T + tt (pair
            // pr, shift tt - nphi + 1).  A translation maps the open hops of L0
            // one to one onto those of every translated word L_ql, with
            // targets in the same orbit, so each open hop of L0 is located
            // once and its Dl images are summed in registers (pair_orbit_sum)
            // before one hash insert.  At half filling only about a third of
            // the candidates are open and a wavefront evaluating them in place
            // would run the lookup for nearly every group of 64, so the open
            // ones are first queued in candidate order (sQ, packed
            // pr << 8 | tt) and evaluated 64 at a time.
            const int ncand = (A.terms & 1) ? npairs * T : 0;
            const int d0 = 64 / T, d1 = 64 % T;  // advance by 64 candidates without division
            int pr = lane / T, tt = lane % T;
            int next = 0, qn = 0;
            for (;;) {
                while (qn < 64 && next < ncand) {
                    bool open = false;
                    if (next + lane < ncand) {
                        const unsigned short pw = sPair[pr];
                        PairHop h;
                        open = pair_hop(nphi, L0, occ[pw & 0xff], occ[pw >> 8], tt - (nphi - 1), h);
                    }
                    const unsigned long long bal = __ballot(open);
                    if (open)
                        sQ[qn + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32),
                                                                __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0u))] =
                            (pr << 8) | tt;
                    qn += __popcll(bal);
                    next += 64;
                    pr += d0;
                    tt += d1;
                    if (tt >= T) {
                        tt -= T;
                        ++pr;
                    }
                }
                if (qn == 0) break;
                __syncthreads();  // queue entries written by other lanes
                const int take = qn < 64 ? qn : 64;
                const int e = lane < take ? sQ[lane] : -1;
                const int carry = lane + 64 < qn ? sQ[lane + 64] : 0;
                __syncthreads();  // every entry read before the carried ones move to the front
                if (lane + 64 < qn) sQ[lane] = carry;
                qn -= take;
                if (e >= 0) {
                    const unsigned short pw = sPair[e >> 8];
                    int j;
                    double re, im;
                    if (pair_orbit_sum(ctx, sLw, sLs, Dl, occ[pw & 0xff], occ[pw >> 8], (e & 0xff) - (nphi - 1), j,
                                       re, im))
                        if (!hash_add<CAP>(keys, vre, vim, j, re, im)) ovf = true;
                }
            }
        }
        for (int ql = 0; ql < Dl; ++ql)
            for (int h = lane; h < A.nhop; h += 64) {
                int j;
                double re, im;
                if (hop_term(ctx, sLw[ql], sLs[ql], Dl, ql, A.hops[h], j, re, im))
                    if (!hash_add<CAP>(keys, vre, vim, j, re, im)) ovf = true;
            }
        __syncthreads();
        // diagonal
        if (lane == 0 && (A.terms & 1)) {
            const double dg = diag_term(g, ctx.spinless != 0, L0, A.Ecn, A.d, A.dv, A.dz);
            if (!hash_add<CAP>(keys, vre, vim, (int)i, dg, 0.0)) ovf = true;
        }
        if (ovf) s_ovf = 1;  // benign race: every writer stores 1
        __syncthreads();
        if (s_ovf) {
            if ((MODE == 0 || MODE == 2) && lane == 0) {
                A.nnz_out[i] = 0;
                const int k = atomicAdd(A.ovf_count, 1);
                A.ovf_rows[k] = (int32_t)i;
                if (MODE == 2) A.stage_ok[i] = 0;
            }
            if (MODE == 1 && lane == 0) atomicAdd(A.err, 1);
            __syncthreads();
            continue;
        }
        if (MODE == 0) {
            int c = 0;
            for (int x = lane; x < CAP; x += 64)
                if (keys[x] >= 0 && hypot(vre[x], vim[x]) > 1e-10) ++c;
            for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
            if (lane == 0) {
                A.nnz_out[i] = c;
                atomicMax(A.row_max, c);
            }
        } else if constexpr (CAP <= 512) {
            // compact the kept (column, value) entries to the front of the
            // hash arrays (lane owns a contiguous chunk of slots, read into
            // registers first), then sort them by column in place
            constexpr int CH = CAP / 64;
            int kk[CH];
            double rr[CH], ii[CH];
            int c = 0;
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                const int x = lane * CH + u;
                rr[u] = vre[x];
                ii[u] = vim[x];
                kk[u] = keys[x] >= 0 && hypot(rr[u], ii[u]) > 1e-10 ? keys[x] : -1;
                c += kk[u] >= 0;
            }
            int pre = c;  // inclusive scan over lanes
            for (int o = 1; o < 64; o <<= 1) {
                const int y = __shfl_up(pre, o, 64);
                if (lane >= o) pre += y;
            }
            const int E = __shfl(pre, 63, 64);
            __syncthreads();  // every slot read before the front is overwritten
            int w = pre - c;
#pragma unroll
            for (int u = 0; u < CH; ++u)
                if (kk[u] >= 0) {
                    keys[w] = kk[u];
                    vre[w] = rr[u];
                    vim[w] = ii[u];
                    ++w;
                }
            int P = 1;
            while (P < E) P <<= 1;
            for (int x = E + lane; x < P; x += 64) keys[x] = 0x7fffffff;  // columns are < 2^31 - 1
            __syncthreads();
            // bitonic sort of keys[0, P) (distinct columns), values moved along
            for (int k = 2; k <= P; k <<= 1)
                for (int jj = k >> 1; jj > 0; jj >>= 1) {
                    for (int x = lane; x < P; x += 64) {
                        const int y = x ^ jj;
                        if (y > x) {
                            const int u = keys[x], v = keys[y];
                            const bool up = (x & k) == 0;
                            if ((u > v) == up) {
                                keys[x] = v;
                                keys[y] = u;
                                const double a = vre[x], b = vim[x];
                                vre[x] = vre[y];
                                vim[x] = vim[y];
                                vre[y] = a;
                                vim[y] = b;
                            }
                        }
                    }
                    __syncthreads();
                }
            if (MODE == 2) {
                // single pass: the sorted row goes to its staging slot (E <= CAP)
                const int64_t base = i * (int64_t)CAP;
                for (int x = lane; x < E; x += 64) {
                    A.stage_col[base + x] = keys[x];
                    A.stage_val[base + x] = make_double2(vre[x], vim[x]);
                }
                if (lane == 0) {
                    A.nnz_out[i] = E;
                    A.stage_ok[i] = 1;
                    atomicMax(A.row_max, E);
                }
            } else {
                const int64_t base = A.rowptr[i];
                const int64_t len = A.rowptr[i + 1] - base;
                if (len != E) {
                    if (lane == 0) atomicAdd(A.err, 1);
                } else {
                    for (int x = lane; x < E; x += 64) {
                        A.col[base + x] = keys[x];
                        A.val[base + x] = make_double2(vre[x], vim[x]);
                    }
                }
            }
        } else {
            // large hash: compact (key, slot) pairs: lane owns a contiguous chunk of slots
            constexpr int CH = CAP / 64;
            int c = 0;
            for (int x = lane * CH; x < lane * CH + CH; ++x)
                if (keys[x] >= 0 && hypot(vre[x], vim[x]) > 1e-10) ++c;
            int pre = c;  // inclusive scan over lanes
            for (int o = 1; o < 64; o <<= 1) {
                const int y = __shfl_up(pre, o, 64);
                if (lane >= o) pre += y;
            }
            const int E = __shfl(pre, 63, 64);
            int w = pre - c;
            for (int x = lane * CH; x < lane * CH + CH; ++x)
                if (keys[x] >= 0 && hypot(vre[x], vim[x]) > 1e-10)
                    sbuf[w++] = ((unsigned long long)(unsigned)keys[x] << 32) | (unsigned)x;
            int P = 1;
            while (P < E) P <<= 1;
            for (int x = E + lane; x < P; x += 64) sbuf[x] = ~0ULL;
            __syncthreads();
            // bitonic sort of sbuf[0, P)
            for (int k = 2; k <= P; k <<= 1)
                for (int jj = k >> 1; jj > 0; jj >>= 1) {
                    for (int x = lane; x < P; x += 64) {
                        const int y = x ^ jj;
                        if (y > x) {
                            const unsigned long long u = sbuf[x], v = sbuf[y];
                            const bool up = (x & k) == 0;
                            if ((u > v) == up) { sbuf[x] = v; sbuf[y] = u; }
                        }
                    }
                    __syncthreads();
                }
            if (MODE == 2) {
                // single pass: the sorted row goes to its staging slot (E <= CAP)
                const int64_t base = i * (int64_t)CAP;
                for (int x = lane; x < E; x += 64) {
                    const unsigned long long u = sbuf[x];
                    const int slot = (int)(u & 0xffffffffu);
                    A.stage_col[base + x] = (int32_t)(u >> 32);
                    A.stage_val[base + x] = make_double2(vre[slot], vim[slot]);
                }
                if (lane == 0) {
                    A.nnz_out[i] = E;
                    A.stage_ok[i] = 1;
                    atomicMax(A.row_max, E);
                }
            } else {
                const int64_t base = A.rowptr[i];
                const int64_t len = A.rowptr[i + 1] - base;
                if (len != E) {
                    if (lane == 0) atomicAdd(A.err, 1);
                } else {
                    for (int x = lane; x < E; x += 64) {
                        const unsigned long long u = sbuf[x];
                        const int slot = (int)(u & 0xffffffffu);
                        A.col[base + x] = (int32_t)(u >> 32);
                        A.val[base + x] = make_double2(vre[slot], vim[slot]);
                    }
                }
            }
        }
        __syncthreads();
    }
}

// Single-pass build: copy every staged row to its CSR position (a wavefront
// per row, entries already sorted); overflow rows are filled by the
// large-hash pass.
template <int CAP>
__global__ void __launch_bounds__(256) k_stage_copy(int64_t n, const int32_t* __restrict__ scol,
                                                    const double2* __restrict__ sval, const uint8_t* __restrict__ ok,
                                                    const int64_t* __restrict__ rowptr, int32_t* __restrict__ col,
                                                    double2* __restrict__ val) {
    const int lane = threadIdx.x & 63;
    const int64_t nw = (int64_t)gridDim.x * 4;
    for (int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); i < n; i += nw) {
        if (!ok[i]) continue;
        const int64_t b = rowptr[i], e = rowptr[i + 1], src = i * (int64_t)CAP;
        for (int64_t x = lane; x < e - b; x += 64) {
            col[b + x] = scol[src + x];
            val[b + x] = sval[src + x];
        }
    }
}

// Row hash capacities: 256 or 512 slots for the first pass (256 when an
// earlier build over the same basis kept at most kSmallRows entries in every
// row: half the LDS, so 50 % more wavefronts per CU), 4096 for the rows that
// overflow it.
constexpr int kSmall = 512, kSmaller = 256, kLarge = 4096, kSmallRows = 192;

size_t table_smem(const Geo& g) {
    const int nphi = g.nphi, T = 2 * nphi - 1, np = g.nel * (g.nel - 1) / 2;
    return sizeof(double) * 2 * nphi * nphi + sizeof(double2) * 2 * T + sizeof(unsigned short) * np + 16;
}

template <class T>
T* upload(DevBuf<T>& buf, const std::vector<T>& v, hipStream_t st) {
    buf.alloc(std::max<size_t>(v.size(), 1));
    if (!v.empty()) BLQH_HIP(hipMemcpyAsync(buf.get(), v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
    return buf.get();
}

template <int CAP, int MODE>
void launch_build(const BuildArgs& A, int64_t rows, size_t smem, hipStream_t st, int ncu) {
    if (rows <= 0) return;
    const int64_t blocks = std::min<int64_t>(rows, (int64_t)ncu * 32);  // rows are claimed by a counter
    BLQH_HIP(hipMemsetAsync(A.row_counter, 0, sizeof(int32_t), st));
    hipLaunchKernelGGL((k_build<CAP, MODE>), dim3((unsigned)blocks), dim3(64), smem, st, A);
    BLQH_HIP(hipGetLastError());
}

template <int MODE>
void launch_first(int cap, const BuildArgs& A, int64_t rows, size_t smem, hipStream_t st, int ncu) {
    if (cap == kSmaller)
        launch_build<kSmaller, MODE>(A, rows, smem, st, ncu);
    else
        launch_build<kSmall, MODE>(A, rows, smem, st, ncu);
}

}  // namespace

// terms = 3: the full Hamiltonian; 1: without the interlayer tunnelling;
// 2: the tunnelling alone at Delta_SAS = 1 (p.delta_sas is ignored), the
// matrix T with H(Delta_SAS) = H(0) + Delta_SAS T (blqh_hamil_combine).
void blqh_hamil_build_impl(const blqh_basis* B, const blqh_params& p_in, blqh_hamil* H, int terms) {
    using namespace blqh;
    BLQH_REQUIRE(terms >= 1 && terms <= 3, BLQH_EINVAL, "terms must be 1, 2 or 3");
    blqh_params p = p_in;
    if (terms == 1) p.delta_sas = 0.0;
    if (terms == 2) p.delta_sas = 1.0;
    H->terms = terms;
    hipStream_t st = stream();
    const int ncu = runtime().num_cus;
    const blqh_sector& s = B->sec;
    BLQH_REQUIRE(p.lx > 0 && p.ly > 0, BLQH_EINVAL, "lx and ly must be positive");
    BLQH_REQUIRE(s.nel <= 64 && s.nel * (s.nel - 1) / 2 < 65536, BLQH_EINVAL, "too many electrons");
    auto t0 = std::chrono::steady_clock::now();
    // BLQH_BUILD_TIMING=1: per-phase wall times of every build on stderr
    static const bool timing = std::getenv("BLQH_BUILD_TIMING") != nullptr;
    double tph[6] = {0, 0, 0, 0, 0, 0};
    auto tmark = t0;
    auto phase = [&](int i) {
        if (!timing) return;
        BLQH_HIP(hipStreamSynchronize(st));
        const auto now = std::chrono::steady_clock::now();
        tph[i] = std::chrono::duration<double, std::milli>(now - tmark).count();
        tmark = now;
    };
    H->basis = B;
    H->par = p;
    H->n = B->n;
    HostTables T = host_tables(s, p);
    phase(0);
    DevBuf<double> dC;
    DevBuf<double2> dFT, dTW;
    DevBuf<HopTerm> dHop;
    BuildArgs A{};
    A.ctx = make_row_ctx(B);
    A.ctx.C = upload(dC, T.C, st);
    A.ctx.FT = upload(dFT, T.FT, st);
    A.ctx.tw = upload(dTW, T.tw, st);
    A.nrows = B->n;
    A.row_list = nullptr;
    A.id = B->id.get();
    A.hops = upload(dHop, T.hops, st);
    A.nhop = (int)T.hops.size();
    A.Ecn = T.Ec * s.nel;
    A.terms = terms;
    A.d = p.d;
    A.dv = p.delta_v;
    A.dz = p.delta_z;
    DevBuf<int32_t> nnz(B->n), ctr(4), ovf_rows(B->n);
    BLQH_HIP(hipMemsetAsync(ctr.get(), 0, 4 * sizeof(int32_t), st));
    A.row_counter = ctr.get();
    A.ovf_count = ctr.get() + 1;
    A.err = ctr.get() + 2;
    A.row_max = ctr.get() + 3;
    const int seen = B->max_row_nnz.load();
    int cap = seen > 0 && seen <= kSmallRows ? kSmaller : kSmall;
    if (const char* e = std::getenv("BLQH_BUILD_CAP")) cap = std::atoi(e) == kSmaller ? kSmaller : kSmall;
    A.ovf_rows = ovf_rows.get();
    A.nnz_out = nnz.get();
    const size_t smem = table_smem(B->geo);
    // One pass when the staging slots fit comfortably (each row hashed once,
    // sorted into a slot of `cap` entries, compacted after the scan); else
    // count, scan, fill.
    size_t free_b = 0, total_b = 0;
    BLQH_HIP(hipMemGetInfo(&free_b, &total_b));
    free_b += dev_cached_bytes();  // cached blocks are released if an allocation needs them
    const size_t stage_bytes = (size_t)B->n * cap * (sizeof(int32_t) + sizeof(double2)) + (size_t)B->n;
    const bool single = stage_bytes <= free_b / 4 && std::getenv("BLQH_BUILD_TWO_PASS") == nullptr;
    DevBuf<int32_t> scol;
    DevBuf<double2> sval;
    DevBuf<uint8_t> sok;
    if (single) {
        scol.alloc((size_t)B->n * cap);
        sval.alloc((size_t)B->n * cap);
        sok.alloc(std::max<int64_t>(B->n, 1));
        A.stage_col = scol.get();
        A.stage_val = sval.get();
        A.stage_ok = sok.get();
        launch_first<2>(cap, A, B->n, smem, st, ncu);
    } else {
        // count
        launch_first<0>(cap, A, B->n, smem, st, ncu);
    }
    int32_t h_ctr[4];
    BLQH_HIP(hipMemcpyAsync(h_ctr, ctr.get(), sizeof(h_ctr), hipMemcpyDeviceToHost, st));
    BLQH_HIP(hipStreamSynchronize(st));
    const int32_t novf = h_ctr[1];
    int32_t row_max = h_ctr[3];
    BuildArgs Ao = A;
    DevBuf<int32_t> ovf2(std::max(novf, 1)), ctr2(4);
    if (novf > 0) {
        // rows that overflowed the small hash: recount with the large one
        BLQH_HIP(hipMemsetAsync(ctr2.get(), 0, 4 * sizeof(int32_t), st));
        Ao.row_list = ovf_rows.get();
        Ao.nrows = novf;
        Ao.row_counter = ctr2.get();
        Ao.ovf_count = ctr2.get() + 1;
        Ao.err = ctr2.get() + 2;
        Ao.row_max = ctr2.get() + 3;
        Ao.ovf_rows = ovf2.get();
        launch_build<kLarge, 0>(Ao, novf, smem, st, ncu);
        BLQH_HIP(hipMemcpyAsync(h_ctr, ctr2.get(), sizeof(h_ctr), hipMemcpyDeviceToHost, st));
        BLQH_HIP(hipStreamSynchronize(st));
        BLQH_REQUIRE(h_ctr[1] == 0, BLQH_ERANGE, "a Hamiltonian row has more than 4096 distinct columns");
        row_max = std::max(row_max, h_ctr[3]);
    }
    // rows of a build with the pair terms size the hash of later builds over this basis
    if (terms & 1) {
        int32_t cur = B->max_row_nnz.load();
        while (row_max > cur && !B->max_row_nnz.compare_exchange_weak(cur, row_max)) {
        }
    }
    phase(1);
    H->rowptr.alloc(B->n + 1);
    exclusive_scan(nnz.get(), H->rowptr.get(), B->n, st);
    copy_sync(&H->nnz, H->rowptr.get() + B->n, 8, hipMemcpyDeviceToHost);
    H->col.alloc(std::max<int64_t>(H->nnz, 1));
    H->val.alloc(std::max<int64_t>(H->nnz, 1));
    phase(2);
    // fill
    A.rowptr = Ao.rowptr = H->rowptr.get();
    A.col = Ao.col = H->col.get();
    A.val = Ao.val = H->val.get();
    BLQH_HIP(hipMemsetAsync(ctr.get(), 0, 4 * sizeof(int32_t), st));
    if (single) {
        const int gs = (int)std::min<int64_t>(std::max<int64_t>(ceil_div(B->n, 4), 1), 65536);
        if (cap == kSmaller)
            hipLaunchKernelGGL(k_stage_copy<kSmaller>, dim3(gs), dim3(256), 0, st, B->n, scol.get(), sval.get(),
                               sok.get(), H->rowptr.get(), H->col.get(), H->val.get());
        else
            hipLaunchKernelGGL(k_stage_copy<kSmall>, dim3(gs), dim3(256), 0, st, B->n, scol.get(), sval.get(),
                               sok.get(), H->rowptr.get(), H->col.get(), H->val.get());
        BLQH_HIP(hipGetLastError());
    } else {
        launch_first<1>(cap, A, B->n, smem, st, ncu);  // overflow rows are skipped (counted in err)
    }
    if (novf > 0) {
        BLQH_HIP(hipMemsetAsync(ctr2.get(), 0, 4 * sizeof(int32_t), st));
        launch_build<kLarge, 1>(Ao, novf, smem, st, ncu);
    }
    int32_t e1[4] = {0, 0, 0, 0}, e2[4] = {0, 0, 0, 0};
    BLQH_HIP(hipMemcpyAsync(e1, ctr.get(), sizeof(e1), hipMemcpyDeviceToHost, st));
    if (novf > 0) BLQH_HIP(hipMemcpyAsync(e2, ctr2.get(), sizeof(e2), hipMemcpyDeviceToHost, st));
    BLQH_HIP(hipStreamSynchronize(st));
    // the small-hash fill pass reports exactly the overflow rows as errors
    BLQH_REQUIRE((single ? e1[2] == 0 : e1[2] == novf) && e2[2] == 0, BLQH_EINTERNAL,
                 "count and fill passes disagree");
    phase(3);
    // the SpMV formats (internal order, panels): now for a full Hamiltonian,
    // on first use for the parts of a merge, which may never need them
    if (terms == 3) build_panels(H);
    phase(4);
    H->seconds_build = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (timing)
        std::fprintf(stderr, "blqh build n=%lld nnz=%lld: tables %.2f count %.2f scan+alloc %.2f fill %.2f panels %.2f "
                     "total %.2f ms\n", (long long)B->n, (long long)H->nnz, tph[0], tph[1], tph[2], tph[3], tph[4],
                     H->seconds_build * 1e3);
}

// ---- H(d, Delta_SAS) = H(d, 0) + Delta_SAS T -------------------------------------
// The interlayer tunnelling moves one electron between the layers, so T's
// entries never share a (row, column) with the layer-conserving Coulomb and
// diagonal terms: the sum is a merge of two sorted rows, with the build's
// drop of entries of modulus <= 1e-10 applied to the scaled T entries.
namespace {

__global__ void k_combine_count(int64_t n, const int64_t* __restrict__ brp, const int64_t* __restrict__ trp,
                                const double2* __restrict__ tval, double dsas, int32_t* __restrict__ cnt) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int c = (int)(brp[i + 1] - brp[i]);
        for (int64_t k = trp[i]; k < trp[i + 1]; ++k) c += hypot(dsas * tval[k].x, dsas * tval[k].y) > 1e-10;
        cnt[i] = c;
    }
}

// One wavefront per row.  The output row is the sorted union of the two
// disjoint rows, so a base entry moves right by the kept T entries of
// smaller column and a kept T entry lands after the kept T entries before it
// and the base entries of smaller column (binary search); every entry is
// placed independently.  A column present in both rows is reported in err.
__global__ void __launch_bounds__(256) k_combine_fill(int64_t n, const int64_t* __restrict__ brp,
                                                      const int32_t* __restrict__ bcol,
                                                      const double2* __restrict__ bval,
                                                      const int64_t* __restrict__ trp,
                                                      const int32_t* __restrict__ tcol,
                                                      const double2* __restrict__ tval, double dsas,
                                                      const int64_t* __restrict__ rp, int32_t* __restrict__ col,
                                                      double2* __restrict__ val, int32_t* __restrict__ err) {
    const int lane = threadIdx.x & 63;
    const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t i = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < n; i += nw) {
        const int64_t a0 = brp[i], na = brp[i + 1] - a0, b0 = trp[i], nb = trp[i + 1] - b0, o0 = rp[i];
        auto kept = [&](int64_t k) { return hypot(dsas * tval[k].x, dsas * tval[k].y) > 1e-10; };
        int64_t placed = 0;  // entries this lane wrote (the row's total is checked below)
        for (int64_t k = lane; k < na; k += 64) {
            const int32_t c = bcol[a0 + k];
            int64_t shift = 0;
            for (int64_t q = 0; q < nb; ++q) shift += (tcol[b0 + q] < c && kept(b0 + q)) ? 1 : 0;
            col[o0 + k + shift] = c;
            val[o0 + k + shift] = bval[a0 + k];
            ++placed;
        }
        for (int64_t q = lane; q < nb; q += 64) {
            if (!kept(b0 + q)) continue;
            int64_t before = 0;
            for (int64_t r = 0; r < q; ++r) before += kept(b0 + r) ? 1 : 0;
            const int32_t c = tcol[b0 + q];
            int64_t lo = 0, hi = na;  // base entries with column < c
            while (lo < hi) {
                const int64_t mid = (lo + hi) >> 1;
                if (bcol[a0 + mid] < c) lo = mid + 1; else hi = mid;
            }
            if (lo < na && bcol[a0 + lo] == c) atomicAdd(err, 1);  // shared entry: not a pure tunnelling term
            col[o0 + before + lo] = c;
            val[o0 + before + lo] = make_double2(dsas * tval[b0 + q].x, dsas * tval[b0 + q].y);
            ++placed;
        }
        for (int o = 32; o > 0; o >>= 1) placed += __shfl_xor(placed, o, 64);
        if (lane == 0 && placed != rp[i + 1] - o0) atomicAdd(err, 1);
    }
}

}  // namespace

void blqh_hamil_combine_impl(const blqh_hamil* base, const blqh_hamil* tun, double dsas, blqh_hamil* H) {
    using namespace blqh;
    BLQH_REQUIRE(base->terms == 1 || (base->terms == 3 && base->par.delta_sas == 0.0), BLQH_EINVAL,
                 "the base must be built without tunnelling (Delta_SAS = 0)");
    BLQH_REQUIRE(tun->terms == 2, BLQH_EINVAL, "the second matrix must be a tunnelling-only build");
    BLQH_REQUIRE(base->basis == tun->basis, BLQH_EINVAL, "the two matrices belong to different bases");
    const blqh_params &a = base->par, &b = tun->par;
    // T depends on d only through the theta_B phases
    BLQH_REQUIRE(a.lx == b.lx && a.ly == b.ly && a.nLL == b.nLL && (a.d == b.d || a.theta_b == 0.0) &&
                     a.theta_b == b.theta_b && a.theta_x == b.theta_x && a.theta_y == b.theta_y,
                 BLQH_EINVAL, "the two matrices were built for different parameters");
    BLQH_REQUIRE(std::isfinite(dsas) && dsas >= 0.0, BLQH_EINVAL, "Delta_SAS must be finite and >= 0");
    auto t0 = std::chrono::steady_clock::now();
    hipStream_t st = stream();
    const int64_t n = base->n;
    H->basis = base->basis;
    H->par = base->par;
    H->par.delta_sas = dsas;
    H->n = n;
    H->terms = 3;
    DevBuf<int32_t> cnt(std::max<int64_t>(n, 1)), err(1);
    BLQH_HIP(hipMemsetAsync(err.get(), 0, sizeof(int32_t), st));
    const int g = (int)std::min<int64_t>(std::max<int64_t>(ceil_div(n, 256), 1), 65536);
    hipLaunchKernelGGL(k_combine_count, dim3(g), dim3(256), 0, st, n, base->rowptr.get(), tun->rowptr.get(),
                       tun->val.get(), dsas, cnt.get());
    H->rowptr.alloc(n + 1);
    exclusive_scan(cnt.get(), H->rowptr.get(), n, st);
    BLQH_HIP(hipMemcpyAsync(&H->nnz, H->rowptr.get() + n, 8, hipMemcpyDeviceToHost, st));
    BLQH_HIP(hipStreamSynchronize(st));
    H->col.alloc(std::max<int64_t>(H->nnz, 1));
    H->val.alloc(std::max<int64_t>(H->nnz, 1));
    const int gw = (int)std::min<int64_t>(std::max<int64_t>(ceil_div(n, 4), 1), 65536);  // four rows per block
    hipLaunchKernelGGL(k_combine_fill, dim3(gw), dim3(256), 0, st, n, base->rowptr.get(), base->col.get(),
                       base->val.get(), tun->rowptr.get(), tun->col.get(), tun->val.get(), dsas, H->rowptr.get(),
                       H->col.get(), H->val.get(), err.get());
    BLQH_HIP(hipGetLastError());
    int32_t herr = 0;
    BLQH_HIP(hipMemcpyAsync(&herr, err.get(), sizeof(int32_t), hipMemcpyDeviceToHost, st));
    BLQH_HIP(hipStreamSynchronize(st));
    BLQH_REQUIRE(herr == 0, BLQH_EINTERNAL, "tunnelling and Coulomb entries overlap");
    build_panels(H);
    H->seconds_build = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}
