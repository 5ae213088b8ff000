// rowterms.hpp -- per-element logic of the basis and Hamiltonian kernels,
// written once for host and device (the CPU test mirror runs the same code).
#pragma once
#include "bits.hpp"

namespace blqh {

// One-body (tunnelling) hop between orbitals n and nt; lo < hi is the range
// whose empty orbitals give the sign (BLQH_without_spin/basis.cpp:187-196).
struct HopTerm {
    int n, nt;
    int lo, hi;
    double coef;  // prefactor including Delta_SAS
    double qre, qim;
    int hasq;
};

struct RowCtx {
    Geo g;
    RankSpec spec;
    int64_t M;      // sector words (lookup-table length)
    int K;          // < 0: no translation reduction (D_l = D_r = 1)
    int spinless;   // operand order of the reference's spinless / spinful expressions
    int twist_sel;  // 1: spinless interlayer pairs, 2: spinful (0,3)/(1,2) pairs
    int mono_odd;   // monolayer with odd nphi: the pair sign counts occupied orbitals (below)
    const int16_t* period;
    const int32_t* lut;
    const uint32_t* cnt;
    const double* C;     // [2][nphi][nphi] Coulomb table V[alpha][|dm|][|t|]
    const double2* FT;   // [2 nphi - 1]: FT[ql - qr + nphi - 1] = exp(2 pi i (ql - qr) K / order)
    const double2* tw;   // [2 nphi - 1] twist phase per t
};

BLQH_HD double2 cmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }

// s mod d from an FP32 reciprocal r of d: the truncated quotient s * r is off
// by at most one for 0 <= s < 2^16, 1 <= d <= 64 and r within 2 ulp of 1/d,
// and one correction step either way makes the remainder exact.
BLQH_HD int mod_by_rcp(int s, int d, float r) {
    const int q = (int)((float)s * r);
    int rem = s - q * d;
    if (rem < 0)
        rem += d;
    else if (rem >= d)
        rem -= d;
    return rem;
}

// s mod d for 0 <= s < 2^16, 1 <= d <= 64.  On the device the quotient comes
// from the hardware's FP32 reciprocal (mod_by_rcp) instead of an integer
// division sequence.
BLQH_HD int mod_small(int s, int d) {
#if defined(__HIP_DEVICE_COMPILE__)
    return mod_by_rcp(s, d, __builtin_amdgcn_rcpf((float)d));
#else
    return s % d;
#endif
}

// Survivor test of the reference's orbit reduction (basis.cpp:82-99): word c
// at position pos(c) of the map's iteration order survives when no other
// orbit member in the sector comes later.
template <class T>
BLQH_HD bool is_survivor(const Geo& g, const RankSpec& s, const T* cnt, const int32_t* posOf, uint64_t c, int32_t p) {
    for (int n = 1; n < g.order(); ++n) {
        int sg;
        const uint64_t c2 = inv_translate(g, c, n, sg);
        if (c2 != c && in_sector(s, c2) && posOf[rank_word(s, cnt, c2)] > p) return false;
    }
    return true;
}

// Lookup-table entry for a sector word y = translate(id[j], q): row j and the
// shift q with inv_translate(y, q) == id[j] as a word.
BLQH_HD int32_t lut_pack(int64_t j, int q) { return (int32_t)((j << 5) | q); }

// Row j of the orbit of sector word Rs, the table's shift q1
// (inv_translate(Rs, q1) == id[j]) and j's period Dr.
BLQH_HD bool locate_word(const RowCtx& c, uint64_t Rs, int& j, int& q1, int& Dr) {
    const int64_t rk = rank_word(c.spec, c.cnt, Rs);
    if (rk < 0 || rk >= c.M) return false;  // defensive: unreachable for sector words
    const int32_t e = c.lut[rk];
    if (e < 0) return false;
    j = e >> 5;
    q1 = e & 31;
    Dr = c.K < 0 ? 1 : c.period[j];
    return true;
}

// Representative of a hop target.  L = translate(id[i], ql) carries momentum
// J + ql*nel, so R0 (same momentum) is first rotated back by ql into the
// sector; the table then gives (j, q1) and the reference's shift is
// qr = (ql + q1) mod D_r, the unique qr < D_r with inv_translate(R0, qr) ==
// id[j] (lanczos_hamiltonian.cpp:147-165).  Its sign is recomputed for qr:
// rotating by the period itself may carry a fermion sign.
BLQH_HD bool locate(const RowCtx& c, uint64_t R0, int ql, int& j, int& qr, int& Dr, int& sr) {
    int s0, q1;
    const uint64_t Rs = ql ? inv_translate(c.g, R0, ql, s0) : R0;
    if (!locate_word(c, Rs, j, q1, Dr)) return false;
    if (c.K < 0) {
        qr = 0;
        sr = 1;
        return true;
    }
    qr = mod_small(ql + q1, Dr);
    (void)inv_translate(c.g, R0, qr, sr);
    return true;
}

// Pair hop n -> n + t, m -> m - t, each orbital shifted within its block
// (layer / species; |t| < nphi) with wrap-around.
struct PairHop {
    int a, b;      // blocks of n and m
    int n0, xm;    // n's orbital and m's target orbital within their blocks
    bool wn, wm;   // the electron wraps around its block
    int nt, mt;    // target orbitals
    uint64_t word; // L with n, m moved to nt, mt
};

// Targets of the pair hop of n and m (both occupied in L) by t; false when
// the hop is blocked (the targets coincide, or another electron of L holds
// one of them).  The compares replace integer divisions.
BLQH_HD bool pair_hop(int nphi, uint64_t L, int n, int m, int t, PairHop& h) {
    h.a = n >= nphi ? (n >= 2 * nphi ? (n >= 3 * nphi ? 3 : 2) : 1) : 0;
    h.b = m >= nphi ? (m >= 2 * nphi ? (m >= 3 * nphi ? 3 : 2) : 1) : 0;
    h.n0 = n - h.a * nphi;
    const int m0 = m - h.b * nphi;
    int xn = h.n0 + t, xm = m0 - t;
    h.wn = xn < 0 || xn >= nphi;
    h.wm = xm < 0 || xm >= nphi;
    xn += xn < 0 ? nphi : (xn >= nphi ? -nphi : 0);
    xm += xm < 0 ? nphi : (xm >= nphi ? -nphi : 0);
    h.xm = xm;
    h.nt = h.a * nphi + xn;
    h.mt = h.b * nphi + xm;
    const uint64_t rest = L ^ ((1ULL << n) | (1ULL << m));
    const uint64_t tgt = (1ULL << h.nt) | (1ULL << h.mt);
    h.word = rest | tgt;
    return h.nt != h.mt && !(rest & tgt);
}

// Two-body element of the open pair hop h of L (pair n < m, shift t) whose
// target lies in the orbit of row j (sector-table shift q1, period Dr).
// BLQH_without_spin/lanczos_hamiltonian.cpp:109-292, BLQH_with_spin/lanczos_hamiltonian.cpp:109-712.
BLQH_HD void pair_value(const RowCtx& c, uint64_t L, int signl, int Dl, int ql, int n, int m, int t, const PairHop& h,
                        int q1, int Dr, double& re, double& im) {
    const int nphi = c.g.nphi;
    const int a = h.a, b = h.b, n0 = h.n0, xm = h.xm, nt = h.nt, mt = h.mt;
    int qr = 0, sr = 1;
    if (c.K >= 0) {
        qr = mod_small(ql + q1, Dr);
        (void)inv_translate(c.g, h.word, qr, sr);
    }
    // The monolayer counts occupied orbitals between n and nt (and m and mt)
    // where sign2 counts empty ones (Monolayer/basis.cpp:169-195).  The counts
    // add up to |nt - n| - 1 and |mt - m| - 1, whose sum is even unless
    // exactly one of the two electrons wraps, and then has the parity of nphi.
    const int sign = sign2(L, n, m, nt, mt) * signl * sr * (c.mono_odd && h.wn != h.wm ? -1 : 1);
    const int sh = c.g.S == 4 ? 1 : 0;  // layer of block a: a / (S / 2)
    const int alpha = (a >> sh) != (b >> sh) ? 1 : 0;
    int dm = xm - n0;
    dm = dm < 0 ? -dm : dm;
    const int at = t < 0 ? -t : t;
    const double Cv = c.C[alpha * nphi * nphi + dm * nphi + at];
    double2 ft = c.FT[ql - qr + nphi - 1];
    const bool twisted = c.twist_sel == 1 ? (a != b) : (c.twist_sel == 2 && ((a == 0 && b == 3) || (a == 1 && b == 2)));
    if (twisted) ft = cmul(ft, c.tw[t + nphi - 1]);
    const double sq = sqrt((double)(Dl * Dr));
    if (c.spinless) {
        const double den = sq * sign;
        re = (ft.x * Cv) / den;
        im = (ft.y * Cv) / den;
    } else {
        const double cs = Cv * sign;
        re = (cs * ft.x) / sq;
        im = (cs * ft.y) / sq;
    }
}

// Two-body term: pair (n < m) occupied in L = translate(id[i], ql), shift t.
// Returns false when the hop is blocked or its target has no representative.
BLQH_HD bool pair_term(const RowCtx& c, uint64_t L, int signl, int Dl, int ql, int n, int m, int t, int& j, double& re,
                       double& im) {
    PairHop h;
    if (!pair_hop(c.g.nphi, L, n, m, t, h)) return false;
    int s0, q1, Dr;
    const uint64_t Rs = ql ? inv_translate(c.g, h.word, ql, s0) : h.word;
    if (!locate_word(c, Rs, j, q1, Dr)) return false;
    pair_value(c, L, signl, Dl, ql, n, m, t, h, q1, Dr, re, im);
    return true;
}

// Orbital x moved by k (0 <= k < nphi) within its block, with wrap-around
// (the translation's action on one orbital).
BLQH_HD int rot_in_block(int nphi, int x, int k) {
    const int a = x >= nphi ? (x >= 2 * nphi ? (x >= 3 * nphi ? 3 : 2) : 1) : 0;
    int y = x - a * nphi + k;
    y -= y >= nphi ? nphi : 0;
    return a * nphi + y;
}

// The pair hop (n0 < m0 of the row word L0 = Ls[0], shift t) and its images
// in the translated words Ls[ql] = translate(L0, ql) (signs sgl[ql]), ql < Dl.
// A translation maps the open hops of L0 one to one onto those of Ls[ql] (the
// image moves the rotated orbitals; when the rotation swaps their order the
// image is the pair (m, n) with shift -t), and maps every hop target onto the
// same orbit, so the target is located once.  Returns the row j and the sum
// of the Dl elements, ql ascending, each computed as pair_term computes the
// image's term in Ls[ql]; false when the hop is blocked or its target has no
// representative.
BLQH_HD bool pair_orbit_sum(const RowCtx& c, const uint64_t* Ls, const int* sgl, int Dl, int n0, int m0, int t, int& j,
                            double& re, double& im) {
    const int nphi = c.g.nphi;
    PairHop h;
    if (!pair_hop(nphi, Ls[0], n0, m0, t, h)) return false;
    int q1, Dr;
    if (!locate_word(c, h.word, j, q1, Dr)) return false;
    re = 0.0;
    im = 0.0;
    for (int ql = 0; ql < Dl; ++ql) {
        const int k = ql * c.g.step;
        int n = rot_in_block(nphi, n0, k), m = rot_in_block(nphi, m0, k), tq = t;
        if (n > m) {
            const int x = n;
            n = m;
            m = x;
            tq = -t;
        }
        if (ql) (void)pair_hop(nphi, Ls[ql], n, m, tq, h);  // open: the image of an open hop
        double r, i;
        pair_value(c, Ls[ql], sgl[ql], Dl, ql, n, m, tq, h, q1, Dr, r, i);
        re += r;
        im += i;
    }
    return true;
}

// One-body tunnelling term.  BLQH_without_spin/lanczos_hamiltonian.cpp:295-322,
// BLQH_with_spin/lanczos_hamiltonian.cpp:718-898.
BLQH_HD bool hop_term(const RowCtx& c, uint64_t L, int signl, int Dl, int ql, const HopTerm& h, int& j, double& re,
                      double& im) {
    const uint64_t mask = (1ULL << h.n) | (1ULL << h.nt);
    const uint64_t kn = L & mask;
    if (kn == 0 || kn == mask) return false;
    const uint64_t R0 = L ^ mask;
    if (c.spec.maskA && popc(R0 & c.spec.maskA) != c.spec.nA) return false;
    if (c.spec.maskB && popc(R0 & c.spec.maskB) != c.spec.nB) return false;
    int qr, Dr, sr;
    if (!locate(c, R0, ql, j, qr, Dr, sr)) return false;
    const int sign = sign1(L, h.lo, h.hi) * signl * sr;
    const int nphi = c.g.nphi;
    const double cs = h.coef * sign;
    const double2 f = c.FT[ql - qr + nphi - 1];
    double2 v = make_double2(cs * f.x, cs * f.y);
    if (h.hasq) v = cmul(v, make_double2(h.qre, h.qim));
    const double sq = sqrt((double)(Dl * Dr));
    re = v.x / sq;
    im = v.y / sq;
    return true;
}

// Diagonal: background Ec*nel, charging -d n_U n_L / nphi, bias, Zeeman.
// BLQH_without_spin/lanczos_hamiltonian.cpp:324-339, BLQH_with_spin/lanczos_hamiltonian.cpp:901-908.
BLQH_HD double diag_term(const Geo& g, bool spinless, uint64_t c, double Ecn, double d, double dv, double dz) {
    const int nphi = g.nphi;
    const uint64_t m = lowbits(nphi);
    if (spinless) {
        const int nu = popc(c & m), nl = popc((c >> nphi) & m);
        return (dv > 0 ? 0.5 * dv * (nu - nl) : 0.0) + Ecn + (-d * nu * (g.nel - nu)) / nphi;
    }
    const int n0 = popc(c & m), n1 = popc((c >> nphi) & m), n2 = popc((c >> 2 * nphi) & m),
              n3 = popc((c >> 3 * nphi) & m);
    return Ecn + (-d * (n0 + n1) * (n2 + n3)) / nphi + (-0.5 * dv) * (n0 + n1 - n2 - n3) +
           (-0.5 * dz) * (n0 + n2 - n1 - n3);
}

// Pair index x (row-major over ia < ib) -> (ia, ib) packed as ia | ib << 8.
BLQH_HD unsigned short pair_decode(int nel, int x) {
    int ia = 0, rem = x;
    while (rem >= nel - 1 - ia) {
        rem -= nel - 1 - ia;
        ++ia;
    }
    return (unsigned short)(ia | ((ia + 1 + rem) << 8));
}

}  // namespace blqh
