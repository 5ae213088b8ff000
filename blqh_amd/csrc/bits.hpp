// bits.hpp -- occupation-word primitives shared by host and device code.
//
// A configuration is a 64-bit word: species block b (b = 0..S-1) holds
// orbitals [b*nphi, (b+1)*nphi).  S = 2: upper layer, lower layer
// (BLQH_without_spin).  S = 4: upper-up, upper-down, lower-up, lower-down
// (BLQH_with_spin/basis.cpp:252-289).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define BLQH_HD __host__ __device__ __forceinline__

namespace blqh {

BLQH_HD int popc(uint64_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __popcll(x);
#else
    return __builtin_popcountll(x);
#endif
}

BLQH_HD uint64_t lowbits(int k) { return k >= 64 ? ~0ULL : ((1ULL << k) - 1ULL); }
// bits [a, b)
BLQH_HD uint64_t range_bits(int a, int b) { return b <= a ? 0ULL : (lowbits(b) & ~lowbits(a)); }

struct Geo {
    int S;         // species blocks (1: monolayer)
    int nphi;      // orbitals per block
    int nel;
    int step = 1;  // orbitals per translation unit: 1 in the bilayer trees,
                   // nphi / gcd(nel, nphi) in the monolayer
    // translation units before the identity
    BLQH_HD int order() const { return nphi / step; }
};

// Magnetic translation by k units (rotate every block left by k * step) with
// fermion sign (-1)^{sum_b (n_b - w_b) w_b}, w_b = electrons of block b that
// wrap.  Reference: BLQH_without_spin/basis.cpp:292-316 (step 1),
// Monolayer/basis.cpp:257-272 (step nphi / gcd).
BLQH_HD uint64_t translate(const Geo& g, uint64_t c, int k, int& sign) {
    if (k == 0) { sign = 1; return c; }
    k *= g.step;
    const uint64_t m = lowbits(g.nphi);
    uint64_t out = 0;
    int ns = 0;
    for (int b = 0; b < g.S; ++b) {
        const uint64_t blk = (c >> (b * g.nphi)) & m;
        const int w = popc(blk >> (g.nphi - k));
        ns += (popc(blk) - w) * w;
        out |= (((blk << k) | (blk >> (g.nphi - k))) & m) << (b * g.nphi);
    }
    sign = (ns & 1) ? -1 : 1;
    return out;
}

// Inverse translation (rotate right by k units).  BLQH_without_spin/basis.cpp:318-340,
// Monolayer/basis.cpp:274-289.
BLQH_HD uint64_t inv_translate(const Geo& g, uint64_t c, int k, int& sign) {
    if (k == 0) { sign = 1; return c; }
    k *= g.step;
    const uint64_t m = lowbits(g.nphi);
    uint64_t out = 0;
    int ns = 0;
    for (int b = 0; b < g.S; ++b) {
        const uint64_t blk = (c >> (b * g.nphi)) & m;
        const int w = popc(blk & lowbits(k));
        ns += (popc(blk) - w) * w;
        out |= (((blk >> k) | (blk << (g.nphi - k))) & m) << (b * g.nphi);
    }
    sign = (ns & 1) ? -1 : 1;
    return out;
}

// Empty orbitals of w strictly between x and y.
BLQH_HD int empties_between(uint64_t w, int x, int y) {
    const int lo = x < y ? x : y, hi = x < y ? y : x;
    if (hi - lo <= 1) return 0;
    return (hi - lo - 1) - popc(w & range_bits(lo + 1, hi));
}

// Two-body hop sign for c -> c - n - m + nt + mt: parity of the empty
// orbitals of c^(n|m) strictly between n,nt and between m,mt, plus one when
// the pair crosses.  BLQH_without_spin/basis.cpp:158-184.
BLQH_HD int sign2(uint64_t c, int n, int m, int nt, int mt) {
    const uint64_t b = c ^ ((1ULL << n) | (1ULL << m));
    int ns = empties_between(b, n, nt) + empties_between(b, m, mt);
    if ((nt > mt && m > n) || (mt > nt && m < n)) ns++;
    return (ns & 1) ? -1 : 1;
}

// One-body hop sign: parity of empty orbitals of c strictly between n and nt,
// counted only when n < nt (the reference loop runs from n+1 to nt-1).
// BLQH_without_spin/basis.cpp:187-196.
BLQH_HD int sign1(uint64_t c, int n, int nt) {
    if (nt <= n + 1) return 1;
    const int e = (nt - n - 1) - popc(c & range_bits(n + 1, nt));
    return (e & 1) ? -1 : 1;
}

// ---------------------------------------------------------------------------
// Combinatorial ranking of a sector.
//
// Sector = words with popcount nel, sum over electrons of (orbital mod nphi)
// == J (mod R; R = nphi with a J filter, 1 without), |w & A| == nA and
// |w & B| == nB (A = upper layer when the layer occupation is fixed,
// B = spin-up orbitals when the spin-up count is fixed; empty masks mean no
// constraint).  cnt[p][k][r][a][b] counts subsets of [0, p) with k elements,
// residual r, a elements in A and b in B.  rank(w) = number of sector words
// numerically below w; words ascend with rank, so unrank(0..M-1) lists the
// sector in ascending order.
// ---------------------------------------------------------------------------
struct RankSpec {
    int nbits, nphi, nel, R, Jt;
    uint64_t maskA, maskB;
    int nA, nB;
    int dK, dR, dA, dB;  // table extents
    int64_t sP, sK, sR, sA;  // strides (b stride is 1)
};

BLQH_HD int64_t rs_index(const RankSpec& s, int p, int k, int r, int a, int b) {
    return p * s.sP + k * s.sK + r * s.sR + a * s.sA + b;
}

// Is w a member of the sector?
BLQH_HD bool in_sector(const RankSpec& s, uint64_t w) {
    if (popc(w) != s.nel) return false;
    if (s.maskA && popc(w & s.maskA) != s.nA) return false;
    if (s.maskB && popc(w & s.maskB) != s.nB) return false;
    if (s.R > 1) {
        // sum of orbital indices mod nphi, block by block
        int js = 0;
        uint64_t x = w;
        while (x) {
#if defined(__HIP_DEVICE_COMPILE__)
            const int p = __ffsll((unsigned long long)x) - 1;
#else
            const int p = __builtin_ctzll(x);
#endif
            js += p % s.nphi;
            x &= x - 1;
        }
        if (js % s.R != s.Jt) return false;
    }
    return true;
}

// Highest set bit of x != 0.
BLQH_HD int top_bit(uint64_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return 63 - __clzll((long long)x);
#else
    return 63 - __builtin_clzll(x);
#endif
}

// rank of a sector word (precondition: in_sector(w)).  cnt is the table
// (fewer than 2^31 entries).  The walk visits the set bits from the top, so
// every word of the sector takes exactly nel steps (no lane of a wavefront
// waits on another's empty orbitals); the table index is 32-bit.
template <class T>
BLQH_HD int64_t rank_word(const RankSpec& s, const T* cnt, uint64_t w) {
    int k = s.nel, r = s.Jt, a = s.nA, b = s.nB;
    const int sP = (int)s.sP, sK = (int)s.sK, sR = (int)s.sR, sA = (int)s.sA, nphi = s.nphi;
    int64_t rk = 0;
    for (; k > 0 && w; --k) {
        const int p = top_bit(w);
        w ^= 1ULL << p;
        rk += (int64_t)cnt[p * sP + k * sK + r * sR + a * sA + b];
        const int pm = p - (p >= nphi ? nphi : 0) - (p >= 2 * nphi ? nphi : 0) - (p >= 3 * nphi ? nphi : 0);
        r -= pm;
        if (r < 0) r += s.R;
        if (s.R == 1) r = 0;
        a -= (int)((s.maskA >> p) & 1ULL);
        b -= (int)((s.maskB >> p) & 1ULL);
    }
    return rk;
}

template <class T>
BLQH_HD uint64_t unrank_word(const RankSpec& s, const T* cnt, int64_t rk) {
    int k = s.nel, r = s.Jt, a = s.nA, b = s.nB;
    uint64_t w = 0;
    for (int p = s.nbits - 1; p >= 0 && k > 0; --p) {
        const int64_t c0 = (int64_t)cnt[rs_index(s, p, k, r, a, b)];
        if (rk < c0) continue;
        rk -= c0;
        w |= 1ULL << p;
        --k;
        r -= p % s.nphi;
        if (r < 0) r += s.R;
        if (s.R == 1) r = 0;
        if ((s.maskA >> p) & 1ULL) --a;
        if ((s.maskB >> p) & 1ULL) --b;
    }
    return w;
}

}  // namespace blqh
