// blqh_oracle.cpp -- CPU restatement of the reference path
//   basis -> hamiltonian (CSR) -> three-pass Lanczos
// of qiongz/BLQH (spinless: BLQH_without_spin/, spinful: BLQH_with_spin/).
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
// product (blqh_amd/) never links or calls it.
//
// Fidelity notes
//  * The basis keeps the reference's representative rule: every J-sector
//    configuration is inserted into a std::unordered_map<unsigned long,long>
//    in the reference's depth-first order, and an orbit member is erased while
//    iterating if another member is still present, so the survivor is the
//    member that comes last in this libstdc++'s iteration order.
//  * Each Hamiltonian row is accumulated column by column in the reference's
//    loop order (ql, pair block, n, m, t, qr, then tunnelling, then the
//    diagonal), with the reference's operand order in every expression, so
//    values match a -ffp-contract=off build of the reference bit for bit.
//  * Lanczos pass 1 and pass 3 follow the reference's vector-op sequences
//    (pass 1 subtracts the complex overlap, pass 3 the real one).
//  * Pass 2 uses a cyclic Jacobi eigensolver written here (the reference calls
//    LAPACK dsyevd), so the Krylov eigenvector may differ in sign and last bits.
//  * Spinless all-density generation uses 64-bit words.  The reference stores
//    the word in an int (BLQH_without_spin/basis.cpp:241), which is identical
//    for 2*nphi <= 31 and corrupts keys at 2*nphi = 32; this oracle keeps the
//    uncorrupted semantics there.
//
// What pins each part (tests/): the basis, translations and signs against the
// reference's own basis.cpp compiled from its sources (oracle/_ref); the
// README's dimensions, energies and amplitudes; the (J, K) Hamiltonian against
// the projection of its K = -1 (plain Fock word) matrix onto truncated Bloch
// vectors, entry by entry (phases, 1/sqrt(D_l D_r), translation signs); the
// spinful pair blocks by the spin-polarised = spinless spectrum and by SU(2)
// multiplets at zero Zeeman energy; Lanczos against LAPACK; the RNG against
// MT19937's published 10000th draw.  The monolayer sectors against the
// Monolayer tree's check data (35 ground energies per particle over three
// electron numbers and the 10 lowest levels of all 108 (J, K) sectors at
// nphi = 18, nel = 6, 6 significant digits) and the same Bloch projection.
// g_plant plants mistakes for the tests that show these checks fail on them.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <random>
#include <unordered_map>
#include <vector>
#include <omp.h>

using cplx = std::complex<double>;

extern "C" {
struct orc_sector {
    int S;       // species: 2 = spinless bilayer, 4 = spinful bilayer
    int nphi;    // orbitals per species
    int nel;     // electrons
    int nel_up;  // electrons in the upper layer, -1 = all densities
    int nel_su;  // spin-up electrons (S = 4), -1 = unconstrained
    int J, K;    // momentum sector, -1 = no filter / no translation reduction
};
struct orc_params {
    double lx, ly;
    int nLL;
    double d, dsas, dv, dz, theta_b, theta_x, theta_y;
};
}

namespace {

inline int pc(uint64_t x) { return __builtin_popcountll(x); }
inline uint64_t lowmask(int k) { return k >= 64 ? ~0ULL : ((1ULL << k) - 1); }

// Magnetic translation by k orbitals on every species block, with the fermion
// sign (-1)^{sum_blocks (n_b - w_b) w_b}, w_b = electrons wrapping around.
// BLQH_without_spin/basis.cpp:292-316, BLQH_with_spin/basis.cpp:252-289.
uint64_t translate(const orc_sector& s, uint64_t c, int k, int& sign) {
    const uint64_t m = lowmask(s.nphi);
    uint64_t out = 0;
    long ns = 0;
    for (int b = 0; b < s.S; ++b) {
        uint64_t blk = (c >> (b * s.nphi)) & m;
        int w = pc(blk & (lowmask(k) << (s.nphi - k)));
        ns += (long)(pc(blk) - w) * w;
        uint64_t r = k == 0 ? blk : (((blk << k) | (blk >> (s.nphi - k))) & m);
        out |= r << (b * s.nphi);
    }
    sign = (ns % 2 == 0) ? 1 : -1;
    return out;
}

// Inverse translation (rotate right by k).  BLQH_without_spin/basis.cpp:318-340,
// BLQH_with_spin/basis.cpp:291-328.
uint64_t inv_translate(const orc_sector& s, uint64_t c, int k, int& sign) {
    const uint64_t m = lowmask(s.nphi);
    uint64_t out = 0;
    long ns = 0;
    for (int b = 0; b < s.S; ++b) {
        uint64_t blk = (c >> (b * s.nphi)) & m;
        int w = pc(blk & lowmask(k));
        ns += (long)(pc(blk) - w) * w;
        uint64_t r = k == 0 ? blk : (((blk >> k) | (blk << (s.nphi - k))) & m);
        out |= r << (b * s.nphi);
    }
    sign = (ns % 2 == 0) ? 1 : -1;
    return out;
}

// Two-body hop sign: parity of EMPTY orbitals of b = c^(n|m) strictly between
// n<->nt and m<->mt, plus one if the pair crosses.  BLQH_without_spin/basis.cpp:158-184.
int sign2(uint64_t c, int n, int m, int nt, int mt) {
    uint64_t b = c ^ ((1ULL << n) + (1ULL << m));
    int ns = 0;
    int kl = nt < n ? nt : n, kr = nt < n ? n : nt;
    for (int k = kl + 1; k < kr; ++k) ns += ((b >> k) & 1ULL) == 0;
    kl = mt < m ? mt : m;
    kr = mt < m ? m : mt;
    for (int k = kl + 1; k < kr; ++k) ns += ((b >> k) & 1ULL) == 0;
    if ((nt > mt && m > n) || (mt > nt && m < n)) ns++;
    return (ns % 2 == 0) ? 1 : -1;
}

// One-body hop sign: parity of empty orbitals of c strictly between n and nt
// (loop is empty when nt <= n+1).  BLQH_without_spin/basis.cpp:187-196.
int sign1(uint64_t c, int n, int nt) {
    int ns = 0;
    for (int i = n + 1; i < nt; ++i) ns += ((c >> i) & 1ULL) == 0;
    return (ns % 2 == 0) ? 1 : -1;
}

struct Basis {
    orc_sector s;
    std::vector<uint64_t> id;
    std::vector<int16_t> period;
    std::unordered_map<uint64_t, long> index;
    long find(uint64_t c) const {
        auto it = index.find(c);
        return it == index.end() ? -1 : it->second;
    }
};

// Depth-first enumeration in the reference's order: the first nel_up
// electrons in the upper layer (bits [0, W)), the rest in [W, 2W), each
// position tuple visited lexicographically; with nel_up < 0 all positions in
// [0, S*nphi).  J filter: sum of (position mod nphi) == J.
// BLQH_without_spin/basis.cpp:199-259, BLQH_with_spin/basis.cpp:240-299.
template <class F>
void enumerate_dfs(const orc_sector& s, F&& emit) {
    const int nbits = s.S * s.nphi;
    const int W = nbits / 2;
    std::vector<int> lo(s.nel), hi(s.nel);
    for (int k = 0; k < s.nel; ++k) {
        if (s.nel_up >= 0) {
            if (k < s.nel_up) { lo[k] = 0; hi[k] = W - (s.nel_up - k - 1); }
            else { lo[k] = W; hi[k] = nbits - (s.nel - k - 1); }
        } else { lo[k] = 0; hi[k] = nbits - (s.nel - k - 1); }
    }
    std::vector<int> pos(s.nel + 1);
    // iterative DFS over position tuples
    int k = 0;
    pos[0] = lo[0];
    if (s.nel == 0) { if (s.J < 0 || s.J == 0) emit(0ULL); return; }
    std::vector<uint64_t> cfg(s.nel + 1, 0);
    std::vector<long> jsum(s.nel + 1, 0);
    while (k >= 0) {
        if (pos[k] >= hi[k]) {  // exhausted this level
            --k;
            if (k >= 0) pos[k]++;
            continue;
        }
        uint64_t c = cfg[k] + (1ULL << pos[k]);
        long js = jsum[k] + pos[k] % s.nphi;
        if (k + 1 == s.nel) {
            if (s.J < 0 || js % s.nphi == s.J) emit(c);
            pos[k]++;
        } else {
            cfg[k + 1] = c;
            jsum[k + 1] = js;
            int start = pos[k] + 1;
            if (s.nel_up >= 0 && k + 1 == s.nel_up) start = W;
            if (start < lo[k + 1]) start = lo[k + 1];
            pos[k + 1] = start;
            ++k;
        }
    }
}

int spin_up_count(const orc_sector& s, uint64_t c) {
    const uint64_t m = lowmask(s.nphi);
    return pc(c & m) + pc((c >> (2 * s.nphi)) & m);
}

// basis::init, BLQH_without_spin/basis.cpp:61-120 and BLQH_with_spin/basis.cpp:74-141.
void build_basis(Basis& B) {
    const orc_sector& s = B.s;
    std::unordered_map<unsigned long, long> set;
    enumerate_dfs(s, [&](uint64_t c) { set[c] = (long)c; });
    int sg;
    if (s.K >= 0) {
        for (auto it = set.begin(); it != set.end();) {
            unsigned long c = it->first;
            bool del = false;
            for (int n = 1; n < s.nphi; ++n) {
                unsigned long cfg = inv_translate(s, c, n, sg);
                if (set.find(cfg) != set.end() && cfg != c) { del = true; break; }
            }
            if (del) it = set.erase(it); else ++it;
        }
    }
    if (s.S == 4 && s.nel_su >= 0) {
        for (auto it = set.begin(); it != set.end();) {
            if (spin_up_count(s, it->first) != s.nel_su) it = set.erase(it); else ++it;
        }
    }
    B.id.clear();
    for (auto& kv : set) B.id.push_back((uint64_t)kv.second);
    std::sort(B.id.begin(), B.id.end());
    B.index.clear();
    for (size_t i = 0; i < B.id.size(); ++i) B.index.insert({B.id[i], (long)i});
    B.period.assign(B.id.size(), (int16_t)s.nphi);
    for (size_t i = 0; i < B.id.size(); ++i)
        for (int n = 1; n < s.nphi; ++n)
            if (inv_translate(s, B.id[i], n, sg) == B.id[i]) { B.period[i] = (int16_t)n; break; }
}

struct Tables {
    std::vector<double> C;  // [alpha][s][|t|]
    std::vector<cplx> FT;   // [ql][qr]
    double Ec;
};

// Coulomb_interaction + init_Coulomb_matrix, BLQH_without_spin/lanczos_hamiltonian.cpp:37-85
// (spinful: BLQH_with_spin/lanczos_hamiltonian.cpp:37-84, theta_x enters as -theta_x).
double coulomb_v(const orc_params& p, int alpha, int qx, int qy) {
    double q = std::sqrt(qx * qx / (p.lx * p.lx) + qy * qy / (p.ly * p.ly)) * 2.0 * M_PI;
    if (alpha == 1)
        return 2.0 * M_PI / q * std::exp(-q * q / 2.0 - q * p.d) * std::pow(1.0 - std::exp(-q * q / 2.0), p.nLL * 2);
    return 2.0 * M_PI / q * std::exp(-q * q / 2.0) * std::pow(1.0 - std::exp(-q * q / 2.0), p.nLL * 2);
}

Tables make_tables(const orc_sector& s, const orc_params& p) {
    Tables T;
    const int nphi = s.nphi;
    const double lx = p.lx, ly = p.ly;
    T.Ec = -2.0;
    for (int i = 0; i < nphi; ++i)
        for (int j = 0; j < nphi; ++j)
            if (!(i == 0 && j == 0)) T.Ec += std::erfc(std::sqrt(M_PI * (i * i * lx / ly + j * j * ly / lx)));
    T.Ec /= std::sqrt(lx * ly);
    T.C.assign(2 * nphi * nphi, 0.0);
    for (int alpha = 0; alpha < 2; ++alpha)
        for (int sx = 0; sx < nphi; ++sx)
            for (int qy = 0; qy < nphi; ++qy) {
                double V = 0;
                for (int qx = -nphi; qx <= nphi; ++qx)
                    if (!(qx == 0 && qy == 0))
                        V += 2.0 * coulomb_v(p, alpha, qx, qy) * std::cos(2.0 * M_PI * sx * qx / nphi) / (2.0 * lx * ly);
                if (alpha == 1) {
                    V = 0;
                    for (int qx = -10 * nphi / (p.d + 0.01); qx < 10 * nphi / (p.d + 0.01); ++qx)
                        if (!(qx == 0 && qy == 0)) {
                            double arg = (s.S == 2) ? (2.0 * M_PI * sx + 2 * p.theta_x) : (2.0 * M_PI * sx - p.theta_x);
                            V += 2.0 * coulomb_v(p, alpha, qx, qy) * std::cos(arg * qx / nphi) / (2.0 * lx * ly);
                        }
                }
                T.C[alpha * nphi * nphi + sx * nphi + qy] = V;
            }
    T.FT.assign(nphi * nphi, 0);
    for (int kl = 0; kl < nphi; ++kl)
        for (int kr = 0; kr < nphi; ++kr)
            T.FT[kl * nphi + kr] = cplx(std::cos(2.0 * M_PI * (kl - kr) * s.K / nphi),
                                        std::sin(2.0 * M_PI * (kl - kr) * s.K / nphi));
    return T;
}

struct RowScratch {
    std::vector<cplx> acc;
    std::vector<char> hit;
    std::vector<long> touched;
};

inline void add_to(RowScratch& R, long j, const cplx& v) {
    if (!R.hit[j]) { R.hit[j] = 1; R.touched.push_back(j); }
    R.acc[j] += v;
}

// Period of R0 under translate (1..nphi-1, else nphi).
inline int period_of(const orc_sector& s, uint64_t r0) {
    int sg;
    for (int D = 1; D < s.nphi; ++D)
        if (translate(s, r0, D, sg) == r0) return D;
    return s.nphi;
}

// Planted mistakes, for the tests that show a check fails on them (0 = none):
// 1 drops the spinful pair block (0,1); 2 reads the interlayer table for the
// intra-layer block (2,3); 3 swaps the FT phase's indices (conjugate phase);
// 4 divides by sqrt(D_l) instead of sqrt(D_l D_r).
int g_plant = 0;

// One Hamiltonian row: peer_set_hamil, BLQH_without_spin/lanczos_hamiltonian.cpp:97-347
// and BLQH_with_spin/lanczos_hamiltonian.cpp:97-908.
void build_row(const Basis& B, const Tables& T, const orc_params& p, long i, RowScratch& R,
               std::vector<long>& cols, std::vector<cplx>& vals) {
    const orc_sector& s = B.s;
    const int nphi = s.nphi, S = s.S, K = s.K;
    const bool spinless = (S == 2);
    int Dl = K < 0 ? 1 : B.period[i];
    // pair blocks in the reference's loop order
    static const int blocks2[][2] = {{0, 0}, {1, 1}, {0, 1}};
    static const int blocks4[][2] = {{0, 0}, {1, 1}, {0, 1}, {2, 2}, {3, 3}, {2, 3}, {0, 2}, {1, 3}, {0, 3}, {1, 2}};
    const int nblk = spinless ? 3 : 10;
    const int (*blocks)[2] = spinless ? blocks2 : blocks4;
    for (int ql = 0; ql < Dl; ++ql) {
        int signl = 1;
        uint64_t L = ql == 0 ? B.id[i] : translate(s, B.id[i], ql, signl);
        for (int bi = 0; bi < nblk; ++bi) {
            const int a = blocks[bi][0], bb = blocks[bi][1];
            const int la = a / (S / 2), lb = bb / (S / 2);
            int alpha = la != lb ? 1 : 0;
            if (!spinless && g_plant == 1 && a == 0 && bb == 1) continue;
            if (!spinless && g_plant == 2 && a == 2 && bb == 3) alpha = 1;
            bool twisted;
            if (spinless) twisted = (a != bb);
            else twisted = (a == 0 && bb == 3) || (a == 1 && bb == 2);
            const int n0 = a * nphi, n1 = (a == bb) ? (a + 1) * nphi - 1 : (a + 1) * nphi;
            for (int n = n0; n < n1; ++n) {
                const int m0 = (a == bb) ? n + 1 : bb * nphi, m1 = (bb + 1) * nphi;
                for (int m = m0; m < m1; ++m) {
                    uint64_t mask = (1ULL << n) + (1ULL << m);
                    if ((L & mask) != mask) continue;
                    uint64_t b = L ^ mask;
                    for (int t = -nphi + 1; t < nphi; ++t) {
                        int nt = a * nphi + ((n - a * nphi + t + nphi) % nphi);
                        int mt = bb * nphi + ((m - bb * nphi - t + nphi) % nphi);
                        int sx = std::abs((mt % nphi) - (n % nphi));
                        uint64_t mask_t = (1ULL << nt) + (1ULL << mt);
                        uint64_t occ_t = mask_t & b;
                        uint64_t r0 = mask_t + b;
                        if (!(occ_t == 0 && nt != mt)) continue;
                        int Dr = K < 0 ? 1 : period_of(s, r0);
                        for (int qr = 0; qr < Dr; ++qr) {
                            int signr = 1;
                            uint64_t rb = qr == 0 ? r0 : inv_translate(s, r0, qr, signr);
                            long j = B.find(rb);
                            if (j < 0) continue;
                            int sign = sign2(L, n, m, nt, mt) * signl * signr;
                            const double Cv = T.C[alpha * nphi * nphi + sx * nphi + std::abs(t)];
                            cplx ft = g_plant == 3 ? T.FT[qr * nphi + ql] : T.FT[ql * nphi + qr];
                            const int Dd = g_plant == 4 ? Dl : Dl * Dr;
                            if (spinless) {
                                if (twisted) ft = ft * cplx(std::cos(2 * p.theta_y * t / nphi), std::sin(2 * p.theta_y * t / nphi));
                                add_to(R, j, ft * Cv / (std::sqrt(Dd) * sign));
                            } else {
                                if (twisted) ft = ft * cplx(std::cos(p.theta_y * t / nphi), std::sin(p.theta_y * t / nphi));
                                add_to(R, j, Cv * sign * ft / std::sqrt(Dd));
                            }
                        }
                    }
                }
            }
        }
        if (p.dsas > 0) {
            struct Hop { int nbeg, nend, off; double coef; int qkind; bool swapsign; };
            // qkind: 0 none, 1 +theta_B*n, 2 -theta_B*nt, 3 +theta_B*(n%nphi), 4 -theta_B*(n%nphi)
            std::vector<Hop> hops;
            if (spinless) hops.push_back({0, nphi, nphi, -0.5, 0, false});
            else {
                hops.push_back({0, nphi, 2 * nphi, -0.25, 1, false});
                hops.push_back({2 * nphi, 3 * nphi, -2 * nphi, -0.25, 2, true});
                hops.push_back({nphi, 2 * nphi, 2 * nphi, -0.25, 3, false});
                hops.push_back({3 * nphi, 4 * nphi, -2 * nphi, -0.25, 4, true});
                hops.push_back({0, nphi, 3 * nphi, -0.5, 0, false});
                hops.push_back({nphi, 2 * nphi, nphi, -0.5, 0, false});
            }
            for (const Hop& h : hops)
                for (int n = h.nbeg; n < h.nend; ++n) {
                    int nt = n + h.off;
                    uint64_t mask = (1ULL << n) + (1ULL << nt);
                    uint64_t Kn = mask & L;
                    if (Kn == mask || Kn == 0) continue;
                    uint64_t Ln = Kn ^ mask;
                    uint64_t r0 = L - Kn + Ln;
                    int Dr = K < 0 ? 1 : period_of(s, r0);
                    for (int qr = 0; qr < Dr; ++qr) {
                        int signr = 1;
                        uint64_t rb = qr == 0 ? r0 : inv_translate(s, r0, qr, signr);
                        long j = B.find(rb);
                        if (j < 0) continue;
                        int sign = (h.swapsign ? sign1(L, nt, n) : sign1(L, n, nt)) * signl * signr;
                        const cplx& ft = T.FT[ql * nphi + qr];
                        if (spinless) {
                            add_to(R, j, -(0.5 * p.dsas * sign * ft / std::sqrt(Dl * Dr)));
                        } else {
                            cplx Q(1.0, 0.0);
                            switch (h.qkind) {
                                case 1: Q = cplx(std::cos(p.theta_b * n * p.d), std::sin(p.theta_b * n * p.d)); break;
                                case 2: Q = cplx(std::cos(p.theta_b * nt * p.d), -std::sin(p.theta_b * nt * p.d)); break;
                                case 3: Q = cplx(std::cos(p.theta_b * (n % nphi) * p.d), std::sin(p.theta_b * (n % nphi) * p.d)); break;
                                case 4: Q = cplx(std::cos(p.theta_b * (n % nphi) * p.d), -std::sin(p.theta_b * (n % nphi) * p.d)); break;
                                default: break;
                            }
                            if (h.qkind) add_to(R, j, h.coef * p.dsas * sign * ft * Q / std::sqrt(Dl * Dr));
                            else add_to(R, j, h.coef * p.dsas * sign * ft / std::sqrt(Dl * Dr));
                        }
                    }
                }
        }
    }
    // diagonal
    const uint64_t c = B.id[i];
    const uint64_t m = lowmask(nphi);
    if (spinless) {
        if (p.dv > 0)
            for (int n = 0; n < 2 * nphi; ++n)
                if ((c >> n) & 1ULL) {
                    if (n < nphi) add_to(R, i, cplx(0.5 * p.dv, 0));
                    else add_to(R, i, cplx(-0.5 * p.dv, 0));
                }
        add_to(R, i, cplx(T.Ec * s.nel, 0));
        int nu = pc(c & m);
        add_to(R, i, cplx(-p.d * nu * (s.nel - nu) / nphi, 0));
    } else {
        int n0 = pc(c & m), n1 = pc((c >> nphi) & m), n2 = pc((c >> 2 * nphi) & m), n3 = pc((c >> 3 * nphi) & m);
        add_to(R, i, cplx(T.Ec * s.nel, 0));
        add_to(R, i, cplx(-p.d * (n0 + n1) * (n2 + n3) / nphi, 0));
        add_to(R, i, cplx(-0.5 * p.dv * (n0 + n1 - n2 - n3), 0));
        add_to(R, i, cplx(-0.5 * p.dz * (n0 + n2 - n1 - n3), 0));
    }
    std::sort(R.touched.begin(), R.touched.end());
    cols.clear();
    vals.clear();
    for (long j : R.touched) {
        if (std::abs(R.acc[j]) > 1e-10) { cols.push_back(j); vals.push_back(R.acc[j]); }
        R.acc[j] = 0;
        R.hit[j] = 0;
    }
    R.touched.clear();
}

struct Csr {
    long n = 0;
    std::vector<int64_t> rowptr;
    std::vector<int64_t> col;
    std::vector<cplx> val;
};

// y = H x, per row in column order (Mat::operator*, BLQH_without_spin/matrix.cpp:244-267)
void spmv(const Csr& H, const cplx* x, cplx* y) {
#pragma omp parallel for schedule(dynamic, 256)
    for (long i = 0; i < H.n; ++i) {
        cplx acc = 0;
        for (int64_t k = H.rowptr[i]; k < H.rowptr[i + 1]; ++k) acc += H.val[k] * x[H.col[k]];
        y[i] = acc;
    }
}

// Seeded complex start vector.  std::mt19937(seed); both parts are
// rng()/rng.max()-0.5 passed to one complex<double>(re, im) constructor; g++
// evaluates the imaginary argument first, so each entry takes its imaginary
// part from the first draw.  Normalised with the running sum of |z|^2.
// BLQH_without_spin/lanczos_hamiltonian.cpp:466-483, matrix.cpp:54-75.
void start_vector(long n, unsigned seed, std::vector<cplx>& v) {
    std::mt19937 rng(seed);
    v.resize(n);
    double nrm = 0;
    for (long i = 0; i < n; ++i) {
        double im = rng() * 1.0 / rng.max() - 0.5;
        double re = rng() * 1.0 / rng.max() - 0.5;
        v[i] = cplx(re, im);
        nrm += std::norm(v[i]);
    }
    nrm = std::sqrt(nrm);
    for (long i = 0; i < n; ++i) v[i] /= nrm;
}

// Symmetric eigensolver (cyclic Jacobi with the classical rotation formulas),
// eigenvalues ascending, eigenvectors in columns of V (row-major V[r*m+c]).
void jacobi_eigh(int m, std::vector<double> A, std::vector<double>& w, std::vector<double>& V) {
    V.assign((size_t)m * m, 0.0);
    for (int i = 0; i < m; ++i) V[(size_t)i * m + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0;
        for (int p = 0; p < m; ++p)
            for (int q = p + 1; q < m; ++q) off += A[(size_t)p * m + q] * A[(size_t)p * m + q];
        if (off < 1e-300) break;
        for (int p = 0; p < m; ++p)
            for (int q = p + 1; q < m; ++q) {
                double apq = A[(size_t)p * m + q];
                if (std::fabs(apq) < 1e-300) continue;
                double app = A[(size_t)p * m + p], aqq = A[(size_t)q * m + q];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
                for (int k = 0; k < m; ++k) {
                    double akp = A[(size_t)k * m + p], akq = A[(size_t)k * m + q];
                    A[(size_t)k * m + p] = c * akp - sn * akq;
                    A[(size_t)k * m + q] = sn * akp + c * akq;
                }
                for (int k = 0; k < m; ++k) {
                    double apk = A[(size_t)p * m + k], aqk = A[(size_t)q * m + k];
                    A[(size_t)p * m + k] = c * apk - sn * aqk;
                    A[(size_t)q * m + k] = sn * apk + c * aqk;
                }
                for (int k = 0; k < m; ++k) {
                    double vkp = V[(size_t)k * m + p], vkq = V[(size_t)k * m + q];
                    V[(size_t)k * m + p] = c * vkp - sn * vkq;
                    V[(size_t)k * m + q] = sn * vkp + c * vkq;
                }
            }
    }
    std::vector<int> idx(m);
    for (int i = 0; i < m; ++i) idx[i] = i;
    std::sort(idx.begin(), idx.end(), [&](int x, int y) { return A[(size_t)x * m + x] < A[(size_t)y * m + y]; });
    w.resize(m);
    std::vector<double> V2((size_t)m * m);
    for (int c2 = 0; c2 < m; ++c2) {
        w[c2] = A[(size_t)idx[c2] * m + idx[c2]];
        for (int r = 0; r < m; ++r) V2[(size_t)r * m + c2] = V[(size_t)r * m + idx[c2]];
    }
    V.swap(V2);
}

}  // namespace

extern "C" {

// ---- basis -----------------------------------------------------------------
void* orc_basis_new(const orc_sector* s) {
    Basis* B = new Basis;
    B->s = *s;
    build_basis(*B);
    return B;
}
// Adopt an existing representative list (ascending ids and periods), e.g. to
// build sample rows of a sector whose hash-order reduction is too slow to
// repeat on the CPU; only the index map is built here.
void* orc_basis_adopt(const orc_sector* s, const uint64_t* id, const int16_t* period, int64_t n) {
    Basis* B = new Basis;
    B->s = *s;
    B->id.assign(id, id + n);
    B->period.assign(period, period + n);
    B->index.reserve(n);
    for (int64_t i = 0; i < n; ++i) B->index.insert({B->id[i], (long)i});
    return B;
}
int64_t orc_basis_size(void* h) { return (int64_t)((Basis*)h)->id.size(); }
void orc_basis_export(void* h, uint64_t* id, int16_t* period) {
    Basis* B = (Basis*)h;
    if (id) std::memcpy(id, B->id.data(), B->id.size() * sizeof(uint64_t));
    if (period) std::memcpy(period, B->period.data(), B->period.size() * sizeof(int16_t));
}
int64_t orc_basis_find(void* h, uint64_t c) { return ((Basis*)h)->find(c); }

// Ground-state observables of a state psi (interleaved re/im) in basis h,
// loop for loop as lhamil::occupatation_number, pseudospin_Sz, pseudospin_Sx
// (BLQH_without_spin/lanczos_hamiltonian.cpp:658-735) and the spinful
// occupatation_number, pseudospin_Sz, Sz, upper_Sz, down_Sz, pseudospin_Sx
// (BLQH_with_spin/lanczos_hamiltonian.cpp:1217-1500).
// out: [0] pseudospin Sz, [1] pseudospin Sx, [2] Sz, [3] upper Sz, [4] lower Sz,
// [5 + a*nphi + j] occupation number of orbital j in layer a, [5 + 2*nphi]
// spin-flip tunnelling (spinful spinflip_tunneling, :1519-1592; spinless 0).
void orc_observables(void* h, const double* psi_in, double* out) {
    const Basis& B = *(const Basis*)h;
    const orc_sector& s = B.s;
    const int nphi = s.nphi, K = s.K;
    const long n = (long)B.id.size();
    const cplx* psi = (const cplx*)psi_in;
    std::vector<cplx> FT((size_t)nphi * nphi);
    for (int a = 0; a < nphi; ++a)
        for (int b = 0; b < nphi; ++b)
            FT[a * nphi + b] = cplx(std::cos(2 * M_PI * (a - b) * K / nphi), std::sin(2 * M_PI * (a - b) * K / nphi));
    const uint64_t low = (1ULL << nphi) - 1;
    auto nel_of = [&](uint64_t c, int sp) { return pc((c >> (sp * nphi)) & low); };
    // occupation numbers
    for (int a = 0; a < 2; ++a)
        for (int j = 0; j < nphi; ++j) {
            const uint64_t mask = a == 0 ? (1ULL << j) : (1ULL << (j + nphi));
            cplx occ = 0;
            for (long i = 0; i < n; ++i) {
                const int D = K < 0 ? 1 : B.period[i];
                for (int q = 0; q < D; ++q) {
                    int sign = 1;
                    const uint64_t lb = q == 0 ? B.id[i] : translate(s, B.id[i], q, sign);
                    if ((mask & lb) == mask) occ += std::conj(psi[i]) * psi[i] * FT[q * nphi + q] / (1.0 * D);
                }
            }
            out[5 + a * nphi + j] = occ.real();
        }
    // pseudospin Sz, Sz, upper / lower Sz
    double up = 0, su = 0, us = 0, ls = 0;
    for (long i = 0; i < n; ++i) {
        const double w = std::norm(psi[i]);
        if (s.S == 2) {
            up += nel_of(B.id[i], 0) * w;
        } else {
            const double n0 = nel_of(B.id[i], 0), n1 = nel_of(B.id[i], 1), n2 = nel_of(B.id[i], 2),
                         n3 = nel_of(B.id[i], 3);
            up += (n0 + n1) * w;
            su += (n0 + n2) * w;
            us += (2.0 * n0 - (n0 + n1)) / (2.0 * (n0 + n1) + 1e-8) * w;
            ls += (2.0 * n2 - (n2 + n3)) / (2.0 * (n2 + n3) + 1e-8) * w;
        }
    }
    out[0] = (2.0 * up - s.nel) / (2.0 * s.nel);
    out[2] = s.S == 4 ? (2.0 * su - s.nel) / (2.0 * s.nel) : 0.0;
    out[3] = s.S == 4 ? us : 0.0;
    out[4] = s.S == 4 ? ls : 0.0;
    // pseudospin Sx: tunnelling pairs, coefficient 1/2 (spinless) or 1/4 per loop (spinful)
    struct Loop { int nbeg, nend, off; double coef; };
    std::vector<Loop> loops;
    if (s.S == 2) loops.push_back({0, nphi, nphi, 0.5});
    else {
        loops = {{0, nphi, 2 * nphi, 0.25}, {nphi, 2 * nphi, 2 * nphi, 0.25},
                 {2 * nphi, 3 * nphi, -2 * nphi, 0.25}, {3 * nphi, 4 * nphi, -2 * nphi, 0.25},
                 {0, nphi, 3 * nphi, 0.25}, {3 * nphi, 4 * nphi, -3 * nphi, 0.25},
                 {nphi, 2 * nphi, nphi, 0.25}, {2 * nphi, 3 * nphi, -nphi, 0.25}};
    }
    auto expect = [&](const std::vector<Loop>& lps) {
        cplx acc = 0;
        for (long i = 0; i < n; ++i) {
            const int Dl = K < 0 ? 1 : B.period[i];
            for (int ql = 0; ql < Dl; ++ql) {
                int signl = 1;
                const uint64_t lb = ql == 0 ? B.id[i] : translate(s, B.id[i], ql, signl);
                for (const Loop& lp : lps)
                    for (int m = lp.nbeg; m < lp.nend; ++m) {
                        const int nt = m + lp.off;
                        const uint64_t mask = (1ULL << m) + (1ULL << nt);
                        const uint64_t Kn = mask & lb;
                        if (Kn == mask || Kn == 0) continue;
                        const uint64_t r0 = lb - Kn + (Kn ^ mask);
                        const int Dr = K < 0 ? 1 : period_of(s, r0);
                        for (int qr = 0; qr < Dr; ++qr) {
                            int signr = 1;
                            const uint64_t rb = qr == 0 ? r0 : inv_translate(s, r0, qr, signr);
                            const long j = B.find(rb);
                            if (j < 0) continue;
                            const int sign = (nt > m ? sign1(lb, m, nt) : sign1(lb, nt, m)) * signl * signr;
                            acc += FT[ql * nphi + qr] * std::conj(psi[i]) * psi[j] *
                                   (sign * lp.coef / std::sqrt(Dl * Dr));
                        }
                    }
            }
        }
        return std::abs(acc) / s.nel;
    };
    out[1] = expect(loops);
    // spin-flip tunnelling: upper-up <-> lower-down and upper-down <-> lower-up, 1/2 each (:1519-1592)
    out[5 + 2 * nphi] = s.S == 4 ? expect({{0, nphi, 3 * nphi, 0.5}, {nphi, 2 * nphi, nphi, 0.5}}) : 0.0;
}
void orc_basis_free(void* h) { delete (Basis*)h; }
void orc_plant(int which) { g_plant = which; }

// Number of J-sector configurations before orbit reduction (the reference's
// basis_set size after generate()).
int64_t orc_sector_configs(const orc_sector* s) {
    int64_t n = 0;
    enumerate_dfs(*s, [&](uint64_t) { ++n; });
    return n;
}
// The same configurations in insertion (depth-first) order.
void orc_sector_configs_export(const orc_sector* s, uint64_t* out) {
    int64_t n = 0;
    enumerate_dfs(*s, [&](uint64_t c) { out[n++] = c; });
}

// ---- bit primitives ----------------------------------------------------------
uint64_t orc_translate(const orc_sector* s, uint64_t c, int k, int* sign) { return translate(*s, c, k, *sign); }
uint64_t orc_inv_translate(const orc_sector* s, uint64_t c, int k, int* sign) { return inv_translate(*s, c, k, *sign); }
int orc_sign2(uint64_t c, int n, int m, int nt, int mt) { return sign2(c, n, m, nt, mt); }
int orc_sign1(uint64_t c, int n, int nt) { return sign1(c, n, nt); }

// ---- tables -------------------------------------------------------------------
void orc_tables(const orc_sector* s, const orc_params* p, double* coulomb, double* ec, double* ft) {
    Tables T = make_tables(*s, *p);
    if (coulomb) std::memcpy(coulomb, T.C.data(), T.C.size() * sizeof(double));
    if (ec) *ec = T.Ec;
    if (ft) std::memcpy(ft, T.FT.data(), T.FT.size() * sizeof(cplx));
}

// ---- Hamiltonian ------------------------------------------------------------
void* orc_hamil_new(void* bh, const orc_params* p, int nthreads) {
    Basis* B = (Basis*)bh;
    Tables T = make_tables(B->s, *p);
    const long n = (long)B->id.size();
    Csr* H = new Csr;
    H->n = n;
    std::vector<std::vector<long>> rc(n);
    std::vector<std::vector<cplx>> rv(n);
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
    {
        RowScratch R;
        R.acc.assign(n, 0);
        R.hit.assign(n, 0);
        std::vector<long> cols;
        std::vector<cplx> vals;
#pragma omp for schedule(dynamic, 16)
        for (long i = 0; i < n; ++i) {
            build_row(*B, T, *p, i, R, cols, vals);
            rc[i] = cols;
            rv[i] = vals;
        }
    }
    H->rowptr.assign(n + 1, 0);
    for (long i = 0; i < n; ++i) H->rowptr[i + 1] = H->rowptr[i] + (int64_t)rc[i].size();
    H->col.resize(H->rowptr[n]);
    H->val.resize(H->rowptr[n]);
    for (long i = 0; i < n; ++i) {
        std::copy(rc[i].begin(), rc[i].end(), H->col.begin() + H->rowptr[i]);
        std::copy(rv[i].begin(), rv[i].end(), H->val.begin() + H->rowptr[i]);
    }
    return H;
}
int64_t orc_hamil_nnz(void* h) { return (int64_t)((Csr*)h)->col.size(); }
int64_t orc_hamil_n(void* h) { return (int64_t)((Csr*)h)->n; }
void orc_hamil_export(void* h, int64_t* rowptr, int64_t* col, double* val) {
    Csr* H = (Csr*)h;
    if (rowptr) std::memcpy(rowptr, H->rowptr.data(), H->rowptr.size() * sizeof(int64_t));
    if (col) std::memcpy(col, H->col.data(), H->col.size() * sizeof(int64_t));
    if (val) std::memcpy(val, H->val.data(), H->val.size() * sizeof(cplx));
}
void orc_hamil_free(void* h) { delete (Csr*)h; }

// Build only the listed rows (for row-sampled parity at sizes where the full
// CPU build is too slow).  rowptr_out has nrows+1 entries; returns total nnz,
// or -(needed) if cap is too small.
int64_t orc_hamil_rows(void* bh, const orc_params* p, const int64_t* rows, int64_t nrows, int nthreads,
                       int64_t* rowptr_out, int64_t* col_out, double* val_out, int64_t cap) {
    Basis* B = (Basis*)bh;
    Tables T = make_tables(B->s, *p);
    const long n = (long)B->id.size();
    std::vector<std::vector<long>> rc(nrows);
    std::vector<std::vector<cplx>> rv(nrows);
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
    {
        RowScratch R;
        R.acc.assign(n, 0);
        R.hit.assign(n, 0);
        std::vector<long> cols;
        std::vector<cplx> vals;
#pragma omp for schedule(dynamic, 1)
        for (int64_t r = 0; r < nrows; ++r) {
            build_row(*B, T, *p, rows[r], R, cols, vals);
            rc[r] = cols;
            rv[r] = vals;
        }
    }
    int64_t tot = 0;
    for (auto& v : rc) tot += (int64_t)v.size();
    if (tot > cap) return -tot;
    rowptr_out[0] = 0;
    for (int64_t r = 0; r < nrows; ++r) {
        rowptr_out[r + 1] = rowptr_out[r] + (int64_t)rc[r].size();
        for (size_t k = 0; k < rc[r].size(); ++k) {
            col_out[rowptr_out[r] + k] = rc[r][k];
            val_out[2 * (rowptr_out[r] + k)] = rv[r][k].real();
            val_out[2 * (rowptr_out[r] + k) + 1] = rv[r][k].imag();
        }
    }
    return tot;
}

// Wrap caller-owned CSR arrays (e.g. a sample) as an oracle matrix.
void* orc_csr_wrap(int64_t n, const int64_t* rowptr, const int64_t* col, const double* val) {
    Csr* H = new Csr;
    H->n = n;
    H->rowptr.assign(rowptr, rowptr + n + 1);
    H->col.assign(col, col + rowptr[n]);
    H->val.resize(rowptr[n]);
    std::memcpy((void*)H->val.data(), val, rowptr[n] * sizeof(cplx));
    return H;
}

void orc_spmv(void* h, const double* x, double* y) { spmv(*(Csr*)h, (const cplx*)x, (cplx*)y); }

// ---- Monolayer (single layer) in the conventions of the Monolayer tree ----
// Basis: Monolayer/basis.cpp:55-115 (init), 169-195 (get_sign), 215-241
// (generate), 257-289 (translate / inv_translate).  Hamiltonian:
// Monolayer/lanczos_hamiltonian.cpp:53-176 (init_Coulomb_matrix, set_hamil).
//  * sector words: nel electrons in nphi orbitals, depth-first (lowest orbital
//    of the first electron slowest), J filter sum(orbitals) % nphi == J;
//  * translations: C = gcd(nel, nphi) of them, each a left rotation by
//    q = nphi / C orbitals, sign (-1)^(w (nel - w)) with w the electrons that
//    wrap; with K >= 0 an orbit member is erased while iterating the
//    std::unordered_map when another (different) member is still present,
//    the survivors sorted ascending; basis_C = the orbit's period;
//  * V(s, ky) = sum over kx in [-(nphi/2), nphi/2] (C integer division),
//    (kx, ky) != 0, of 2 pi / (q + 1e-30) exp(-q^2/2) (1 - exp(-q^2/2))^(2 nLL)
//    cos(2 pi s kx / nphi) / (2 lx ly);
//  * pair (n < m) occupied in the k1-translated left word, t in
//    [-(nphi/2), nphi/2], nt = n + t, mt = m - t wrapped once, s = |n - mt|;
//    both targets empty: for every k2 below the target's period, the
//    k2-inverse-translated target, when a basis row, receives
//    2 V(s, |t|) sign FT, FT = exp(2 pi i K (k1 - k2) / C) / sqrt(C_l C_r);
//  * sign: (-1)^(occupied orbitals strictly between n and nt, and between m
//    and mt, in the word without n and m), times -1 when the pair crosses
//    (nt > mt and m > n, or mt > nt and m < n), times both translation signs;
//  * diagonal E_cl * nel, E_cl = (-2 + sum_{(i,j) in [0,nphi)^2 \ 0}
//    I(pi (i^2 lx/ly + j^2 ly/lx))) / sqrt(lx ly) with
//    I(z) = int_1^inf t^-1/2 e^{-zt} dt = sqrt(pi/z) erfc(sqrt(z)) (the tree
//    integrates it numerically with GSL);
//  * entries with |h| > 1e-10 kept, columns ascending.
// A target with nt == mt has nel - 1 electrons in the tree (its mask adds the
// same bit twice) and never matches a row; it is skipped here.
namespace {
struct MonoBasis {
    int nphi = 0, nel = 0, J = -1, K = -1, C = 1, q = 1;
    std::vector<uint64_t> id;
    std::vector<int> per;
    std::unordered_map<unsigned long, long> row;

    uint64_t mask() const { return (1ull << nphi) - 1; }
    uint64_t fwd(uint64_t c, int k, int& sign) const {
        const int r = k * q;
        c &= mask();
        if (r == 0 || r == nphi) { sign = 1; return c; }
        const int w = __builtin_popcountll(c >> (nphi - r));
        sign = ((w * (nel - w)) & 1) ? -1 : 1;
        return ((c << r) | (c >> (nphi - r))) & mask();
    }
    uint64_t bwd(uint64_t c, int k, int& sign) const {
        const int r = k * q;
        c &= mask();
        if (r == 0 || r == nphi) { sign = 1; return c; }
        const int w = __builtin_popcountll(c & ((1ull << r) - 1));
        sign = ((w * (nel - w)) & 1) ? -1 : 1;
        return ((c >> r) | (c << (nphi - r))) & mask();
    }
};

void mono_words(int nphi, int nel, int J, int k, int from, uint64_t c, int jsum,
                std::unordered_map<unsigned long, long>& set) {
    for (int i = from; i < nphi - (nel - k - 1); ++i) {
        const uint64_t w = c | (1ull << i);
        if (k + 1 < nel)
            mono_words(nphi, nel, J, k + 1, i + 1, w, jsum + i, set);
        else if (J < 0 || (jsum + i) % nphi == J)
            set[(unsigned long)w] = (long)w;
    }
}

void mono_basis(int nphi, int nel, int J, int K, MonoBasis& B) {
    B.nphi = nphi;
    B.nel = nel;
    B.J = J;
    B.K = K;
    B.C = 1;
    for (int i = 1; i <= nel; ++i)
        if (nel % i == 0 && nphi % i == 0) B.C = i;
    B.q = nphi / B.C;
    std::unordered_map<unsigned long, long> set;
    mono_words(nphi, nel, J, 0, 0, 0, 0, set);
    if (K >= 0) {
        for (auto it = set.begin(); it != set.end();) {
            const uint64_t c = it->first;
            bool gone = false;
            for (int k = 1; k < B.C && !gone; ++k) {
                int sg;
                const uint64_t y = B.bwd(c, k, sg);
                gone = y != c && set.find((unsigned long)y) != set.end();
            }
            if (gone)
                it = set.erase(it);
            else
                ++it;
        }
    }
    B.id.clear();
    for (const auto& kv : set) B.id.push_back((uint64_t)kv.second);
    std::sort(B.id.begin(), B.id.end());
    B.row.clear();
    for (size_t i = 0; i < B.id.size(); ++i) B.row[(unsigned long)B.id[i]] = (long)i;
    B.per.assign(B.id.size(), B.C);
    for (size_t i = 0; i < B.id.size(); ++i)
        for (int k = 1; k < B.C; ++k) {
            int sg;
            if (B.fwd(B.id[i], k, sg) == B.id[i]) { B.per[i] = k; break; }
        }
}

double mono_tables(int nphi, double lx, double ly, int nLL, std::vector<double>& Cm) {
    auto V = [&](int kx, int ky) {
        const double q = std::sqrt(kx * kx / (lx * lx) + ky * ky / (ly * ly)) * 2.0 * M_PI;
        return 2.0 * M_PI / (q + 1e-30) * std::exp(-q * q / 2.0) * std::pow(1.0 - std::exp(-q * q / 2.0), nLL * 2);
    };
    Cm.assign(nphi * nphi, 0.0);
    for (int s = 0; s < nphi; ++s)
        for (int ky = 0; ky < nphi; ++ky) {
            double v = 0;
            for (int kx = -nphi / 2; kx <= nphi / 2; ++kx)
                if (!(ky == 0 && kx == 0)) v += V(kx, ky) * std::cos(2.0 * M_PI * s * kx / nphi) / (2.0 * lx * ly);
            Cm[s * nphi + ky] = v;
        }
    double ecl = -2.0;
    for (int i = 0; i < nphi; ++i)
        for (int j = 0; j < nphi; ++j)
            if (!(i == 0 && j == 0)) {
                const double z = (i * i * lx / ly + j * j * ly / lx) * M_PI;
                ecl += std::sqrt(M_PI / z) * std::erfc(std::sqrt(z));
            }
    return ecl / std::sqrt(lx * ly);
}

// (-1)^(occupied orbitals of the word without n and m strictly between n and
// nt and between m and mt), times -1 for a crossing pair.
int mono_sign(uint64_t c, int n, int m, int nt, int mt) {
    const uint64_t b = c ^ ((1ull << n) | (1ull << m));
    const int occ = g_plant == 5 ? 0 : 1;  // plant 5: count empty orbitals (the bilayer trees' rule)
    int ns = 0;
    for (int k = std::min(n, nt) + 1; k < std::max(n, nt); ++k) ns += (int)((b >> k) & 1) == occ;
    for (int k = std::min(m, mt) + 1; k < std::max(m, mt); ++k) ns += (int)((b >> k) & 1) == occ;
    if ((nt > mt && m > n) || (mt > nt && m < n)) ++ns;
    return (ns & 1) ? -1 : 1;
}
}  // namespace

// Basis of the (J, K) sector (K < 0: no translation reduction): returns n and
// writes ids / periods when n <= cap (else -n).
int64_t orc_monolayer_basis(int nphi, int nel, int J, int K, uint64_t* id, int16_t* period, int64_t cap) {
    MonoBasis B;
    mono_basis(nphi, nel, J, K, B);
    const int64_t n = (int64_t)B.id.size();
    if (n > cap) return -n;
    for (int64_t i = 0; i < n; ++i) {
        id[i] = B.id[i];
        period[i] = (int16_t)B.per[i];
    }
    return n;
}

void* orc_monolayer_new(int nphi, int nel, int J, int K, double lx, double ly, int nLL, double* ecl_out) {
    MonoBasis B;
    mono_basis(nphi, nel, J, K, B);
    std::vector<double> Cm;
    const double ecl = mono_tables(nphi, lx, ly, nLL, Cm);
    if (ecl_out) *ecl_out = ecl;
    const long n = (long)B.id.size();
    Csr* H = new Csr;
    H->n = n;
    H->rowptr.assign(n + 1, 0);
    std::vector<cplx> row(n, cplx(0, 0));
    std::vector<char> hit(n, 0);
    std::vector<long> touched;
    for (long i = 0; i < n; ++i) {
        touched.clear();
        auto add = [&](long j, cplx v) {
            if (!hit[j]) { hit[j] = 1; touched.push_back(j); }
            row[j] += v;
        };
        const int Cl = K < 0 ? 1 : B.per[i];
        for (int k1 = 0; k1 < Cl; ++k1) {
            int signl = 1;
            const uint64_t L = k1 == 0 ? B.id[i] : B.fwd(B.id[i], k1, signl);
            for (int a = 0; a < nphi - 1; ++a)
                for (int b = a + 1; b < nphi; ++b) {
                    const uint64_t pm = (1ull << a) | (1ull << b);
                    if ((L & pm) != pm) continue;
                    const uint64_t rest = L ^ pm;
                    for (int t = -nphi / 2; t <= nphi / 2; ++t) {
                        int nt = a + t, mt = b - t;
                        nt += nt >= nphi ? -nphi : (nt < 0 ? nphi : 0);
                        mt += mt >= nphi ? -nphi : (mt < 0 ? nphi : 0);
                        if (nt == mt) continue;
                        const int sx = std::abs(a - mt);
                        const uint64_t tg = (1ull << nt) | (1ull << mt);
                        if (tg & rest) continue;
                        const uint64_t R = rest | tg;
                        int Cr = 1;
                        if (K >= 0) {
                            Cr = B.C;
                            for (int k2 = 1; k2 <= B.C; ++k2) {
                                int sg;
                                if (B.fwd(R, k2, sg) == R) { Cr = k2; break; }
                            }
                        }
                        for (int k2 = 0; k2 < Cr; ++k2) {
                            int signr = 1;
                            const uint64_t Rb = k2 == 0 ? R : B.bwd(R, k2, signr);
                            auto it = B.row.find((unsigned long)Rb);
                            if (it == B.row.end()) continue;
                            const int sign = mono_sign(L, a, b, nt, mt) * signl * signr;
                            const double ang = 2.0 * M_PI * K * (g_plant == 6 ? k2 - k1 : k1 - k2) / B.C;
                            const cplx ft = cplx(std::cos(ang), std::sin(ang)) / std::sqrt((double)(Cl * Cr));
                            add(it->second, 2.0 * Cm[sx * nphi + std::abs(t)] * sign * ft);
                        }
                    }
                }
        }
        add(i, cplx(ecl * nel, 0.0));
        std::sort(touched.begin(), touched.end());
        for (long j : touched) {
            if (std::abs(row[j]) > 1e-10) {
                H->col.push_back(j);
                H->val.push_back(row[j]);
            }
            row[j] = cplx(0, 0);
            hit[j] = 0;
        }
        H->rowptr[i + 1] = (int64_t)H->col.size();
    }
    return H;
}

// The same row loop over caller arrays with 32-bit columns (no copy of a
// large matrix); nthreads <= 0: OpenMP's default.
void orc_spmv32(int64_t n, const int64_t* rowptr, const int32_t* col, const double* val, const double* x, double* y,
                int nthreads) {
    const cplx* v = (const cplx*)val;
    const cplx* xx = (const cplx*)x;
    cplx* yy = (cplx*)y;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads)
    for (long i = 0; i < n; ++i) {
        cplx acc = 0;
        for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) acc += v[k] * xx[col[k]];
        yy[i] = acc;
    }
}

void orc_start_vector(int64_t n, unsigned seed, double* out) {
    std::vector<cplx> v;
    start_vector(n, seed, v);
    std::memcpy(out, v.data(), n * sizeof(cplx));
}

void orc_tridiag_eigh(int m, const double* diag, const double* off, double* w, double* V) {
    std::vector<double> A((size_t)m * m, 0.0);
    for (int i = 0; i < m; ++i) A[(size_t)i * m + i] = diag[i];
    for (int i = 0; i + 1 < m; ++i) A[(size_t)i * m + i + 1] = A[(size_t)(i + 1) * m + i] = off[i];
    std::vector<double> ww, VV;
    jacobi_eigh(m, A, ww, VV);
    std::memcpy(w, ww.data(), m * sizeof(double));
    if (V) std::memcpy(V, VV.data(), (size_t)m * m * sizeof(double));
}

// The reference's three-pass Lanczos with exactly `lambda` steps:
// coeff_explicit_update (pass 1), diag (pass 2), eigenstates_reconstruction
// (pass 3) and ground_state_energy, BLQH_without_spin/lanczos_hamiltonian.cpp:453-656.
// alpha[lambda], beta[lambda+1], eig[lambda], psir[2n] (complex), returns
// Re<psir|H psir>.
double orc_lanczos(void* h, int lambda, unsigned seed, double* alpha, double* beta, double* eig, double* psir_out) {
    const Csr& H = *(Csr*)h;
    const long n = H.n;
    std::vector<double> norm(lambda + 1, 0.0), overlap(lambda, 0.0);
    std::vector<cplx> p0, p1(n), p2(n), pt(n);
    start_vector(n, seed, p0);
    // pass 1
    spmv(H, p0.data(), p1.data());
    cplx of = 0;
    for (long i = 0; i < n; ++i) of += std::conj(p1[i]) * p0[i];
    overlap[0] = of.real();
    for (long i = 0; i < n; ++i) p1[i] -= p0[i] * of;
    double nf = 0;
    for (long i = 0; i < n; ++i) nf += std::norm(p1[i]);
    nf = std::sqrt(nf);
    for (long i = 0; i < n; ++i) p1[i] /= nf;
    norm[0] = 1;
    norm[1] = nf;
    for (int j = 1; j < lambda; ++j) {
        spmv(H, p1.data(), p2.data());
        of = 0;
        for (long i = 0; i < n; ++i) of += std::conj(p1[i]) * p2[i];
        overlap[j] = of.real();
        for (long i = 0; i < n; ++i) pt[i] = p0[i] * norm[j];
        for (long i = 0; i < n; ++i) pt[i] += p1[i] * of;
        for (long i = 0; i < n; ++i) p2[i] -= pt[i];
        nf = 0;
        for (long i = 0; i < n; ++i) nf += std::norm(p2[i]);
        nf = std::sqrt(nf);
        for (long i = 0; i < n; ++i) p2[i] /= nf;
        norm[j + 1] = nf;
        std::swap(p0, p1);
        std::swap(p1, p2);
    }
    // pass 2
    const int l = lambda;
    std::vector<double> A((size_t)l * l, 0.0);
    for (int i = 0; i < l - 1; ++i) {
        A[(size_t)i * l + i + 1] = norm[i + 1];
        A[(size_t)(i + 1) * l + i] = norm[i + 1];
        A[(size_t)i * l + i] = overlap[i];
    }
    A[(size_t)(l - 1) * l + l - 1] = overlap[l - 1];
    std::vector<double> w, V;
    jacobi_eigh(l, A, w, V);
    std::vector<double> psi0(l);
    for (int i = 0; i < l; ++i) psi0[i] = V[(size_t)i * l + 0];
    // pass 3
    std::vector<cplx> q0, q1(n), q2(n), psir(n, 0);
    start_vector(n, seed, q0);
    spmv(H, q0.data(), q1.data());
    for (long i = 0; i < n; ++i) q1[i] -= q0[i] * overlap[0];
    for (long i = 0; i < n; ++i) q1[i] /= norm[1];
    for (long i = 0; i < n; ++i) psir[i] += psi0[0] * std::conj(q0[i]) + psi0[1] * std::conj(q1[i]);
    for (int ll = 2; ll < l; ++ll) {
        spmv(H, q1.data(), q2.data());
        for (long i = 0; i < n; ++i) q2[i] -= q1[i] * overlap[ll - 1] + q0[i] * norm[ll - 1];
        for (long i = 0; i < n; ++i) q2[i] /= norm[ll];
        for (long i = 0; i < n; ++i) psir[i] += psi0[ll] * std::conj(q2[i]);
        std::swap(q0, q1);
        std::swap(q1, q2);
    }
    double Nrm = 0;
    for (long i = 0; i < n; ++i) Nrm += std::norm(psir[i]);
    Nrm = std::sqrt(Nrm);
    for (long i = 0; i < n; ++i) psir[i] /= Nrm;
    std::vector<cplx> hp(n);
    spmv(H, psir.data(), hp.data());
    cplx e = 0;
    for (long i = 0; i < n; ++i) e += std::conj(psir[i]) * hp[i];
    if (alpha) std::memcpy(alpha, overlap.data(), l * sizeof(double));
    if (beta) std::memcpy(beta, norm.data(), (l + 1) * sizeof(double));
    if (eig) std::memcpy(eig, w.data(), l * sizeof(double));
    if (psir_out) std::memcpy(psir_out, psir.data(), n * sizeof(cplx));
    return e.real();
}

}  // extern "C"
