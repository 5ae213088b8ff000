// hostcheck.cpp -- CPU mirror of the blqh device pipeline (test support).
//
// Runs the product's own per-element code (bits.hpp / rowterms.hpp: ranking,
// translations, signs, survivor test, lookup packing, pair / hop / diagonal
// terms) and its host steps (rank tables, hash iteration order, tables) on the
// CPU, with plain loops in place of the kernels' scheduling, so the CPU test
// suite can compare the GPU path's arithmetic with the oracle without a GPU.
// The GPU-only parts (LDS hash, bitonic sort, reductions) are covered by the
// GPU tests.
#include <algorithm>
#include <cstring>
#include <map>
#include <vector>

#include "../../blqh_amd/csrc/engine.hpp"
#include "../../blqh_amd/csrc/xcdplan.hpp"
#include "../../blqh_amd/csrc/ritz.hpp"

using namespace blqh;

namespace {
struct HostBasis {
    blqh_sector s;
    Geo g;
    RankSpec spec;
    std::vector<uint32_t> cnt;
    int64_t M = 0;
    std::vector<uint64_t> id;
    std::vector<int16_t> period;
    std::vector<int32_t> lut;
};

void build(const blqh_sector& s, HostBasis& B) {
    B.s = s;
    B.g = make_geo(s);
    RankSpec smap = make_rank_spec(s, false);
    int64_t Mmap = 0;
    std::vector<uint32_t> tmap = make_rank_table(smap, &Mmap);
    std::vector<int32_t> posOf(Mmap, 0);
    const bool reduce = s.K >= 0;
    if (reduce) {
        std::vector<uint64_t> order = hash_iteration_order(s);
        for (size_t p = 0; p < order.size(); ++p) posOf[rank_word(smap, tmap.data(), order[p])] = (int32_t)p;
    }
    const uint64_t spin_mask = lowbits(s.nphi) | (lowbits(s.nphi) << (2 * s.nphi));
    B.id.clear();
    for (int64_t r = 0; r < Mmap; ++r) {
        const uint64_t c = unrank_word(smap, tmap.data(), r);
        bool keep = !reduce || is_survivor(B.g, smap, tmap.data(), posOf.data(), c, posOf[r]);
        if (keep && s.nspecies == 4 && s.nel_su >= 0 && popc(c & spin_mask) != s.nel_su) keep = false;
        if (keep) B.id.push_back(c);
    }
    const int64_t n = (int64_t)B.id.size();
    B.period.assign(n, (int16_t)B.g.order());
    for (int64_t j = 0; j < n; ++j)
        for (int k = 1; k < B.g.order(); ++k) {
            int sg;
            if (inv_translate(B.g, B.id[j], k, sg) == B.id[j]) { B.period[j] = (int16_t)k; break; }
        }
    B.spec = make_rank_spec(s, true);
    B.cnt = make_rank_table(B.spec, &B.M);
    B.lut.assign(B.M, -1);
    for (int64_t j = 0; j < n; ++j) {
        const int D = reduce ? B.period[j] : 1;
        for (int q = 0; q < D; ++q) {
            int sg;
            const uint64_t y = translate(B.g, B.id[j], q, sg);
            if (!in_sector(B.spec, y)) continue;
            B.lut[rank_word(B.spec, B.cnt.data(), y)] = lut_pack(j, q);
        }
    }
}
// The device's level-by-level computation of the map's iteration order
// (basis.hip k_dfs_keys, k_hl_bucket, k_hl_members, k_hl_pos), step by step
// with sequential loops; returns the words in list order.
std::vector<uint64_t> level_order(const blqh_sector& s) {
    const RankSpec smap = make_rank_spec(s, false);
    int64_t M = 0;
    (void)make_rank_table(smap, &M);
    const RankSpec srev = reversed_rank_spec(s, smap);
    int64_t Mrev = 0;
    const std::vector<uint32_t> trev = make_rank_table(srev, &Mrev);
    if (Mrev != M) return {};
    std::vector<uint64_t> key(M);
    for (int64_t i = 0; i < M; ++i)
        key[i] = __builtin_bitreverse64(unrank_word(srev, trev.data(), M - 1 - i)) >> (64 - smap.nbits);
    const auto lv = hash_levels(M);
    std::vector<uint32_t> tau(M), tau2(M), bucket(M), members(M), gmin;
    std::vector<int32_t> gcount, first_sz(M);
    std::vector<int64_t> goff, later(M + 1);
    for (size_t r = 0; r < lv.size(); ++r) {
        const int64_t st = lv[r].first, en = r + 1 < lv.size() ? lv[r + 1].first : M;
        const uint64_t m = lv[r].second;
        if (en == 0) continue;
        gcount.assign(m, 0);
        gmin.assign(m, 0xFFFFFFFFu);
        for (int64_t i = 0; i < en; ++i) {
            if (i >= st) tau[i] = (uint32_t)i;
            bucket[i] = (uint32_t)(key[i] % m);
            ++gcount[bucket[i]];
            gmin[bucket[i]] = std::min(gmin[bucket[i]], tau[i]);
        }
        goff.assign(m + 1, 0);
        for (uint64_t b = 0; b < m; ++b) goff[b + 1] = goff[b] + gcount[b];
        for (int64_t i = 0; i < en; ++i) {
            const uint32_t b = bucket[i];
            members[goff[b] + --gcount[b]] = (uint32_t)i;
            first_sz[en - 1 - tau[i]] = tau[i] == gmin[b] ? (int32_t)(goff[b + 1] - goff[b]) : 0;
        }
        later[0] = 0;
        for (int64_t u = 0; u < en; ++u) later[u + 1] = later[u] + first_sz[u];
        for (int64_t i = 0; i < en; ++i) {
            const uint32_t b = bucket[i], t = tau[i];
            int64_t p = later[en - 1 - gmin[b]];
            for (int64_t k = goff[b]; k < goff[b + 1]; ++k) p += tau[members[k]] > t ? 1 : 0;
            tau2[i] = (uint32_t)p;
        }
        tau.swap(tau2);
    }
    std::vector<uint64_t> out(M);
    for (int64_t i = 0; i < M; ++i) out[tau[i]] = key[i];
    return out;
}
}  // namespace

extern "C" {

// Basis through the mirror: returns n (or -n when cap is too small).
int64_t hc_basis(const blqh_sector* s, uint64_t* id, int16_t* period, int64_t cap) {
    HostBasis B;
    build(*s, B);
    const int64_t n = (int64_t)B.id.size();
    if (n > cap) return -n;
    std::memcpy(id, B.id.data(), n * 8);
    std::memcpy(period, B.period.data(), n * 2);
    return n;
}

// Rank table checks: unrank(0..M-1) into words (ascending), M returned.
int64_t hc_sector_words(const blqh_sector* s, int with_spin, uint64_t* out, int64_t cap) {
    RankSpec sp = make_rank_spec(*s, with_spin != 0);
    int64_t M = 0;
    std::vector<uint32_t> t = make_rank_table(sp, &M);
    if (M > cap) return -M;
    for (int64_t r = 0; r < M; ++r) {
        out[r] = unrank_word(sp, t.data(), r);
        if (!in_sector(sp, out[r]) || rank_word(sp, t.data(), out[r]) != r) return -1 - r - M;  // inconsistency
    }
    return M;
}

uint64_t hc_translate(int S, int nphi, int nel, uint64_t c, int k, int* sign) {
    return translate(Geo{S, nphi, nel}, c, k, *sign);
}
uint64_t hc_inv_translate(int S, int nphi, int nel, uint64_t c, int k, int* sign) {
    return inv_translate(Geo{S, nphi, nel}, c, k, *sign);
}
int hc_sign2(uint64_t c, int n, int m, int nt, int mt) { return sign2(c, n, m, nt, mt); }

// Per-XCD work lists of the batched K1 for P members (member p owns items
// item0[p] .. item0[p+1]); items receives 2 * item0[P] entries, seg 27.
void hc_xcd_lists(int P, const int64_t* item0, const int64_t* nnz, const int32_t* npan, const int32_t* real,
                  int32_t* items, int32_t* seg) {
    std::vector<XcdMember> mem(P);
    for (int p = 0; p < P; ++p) mem[p] = {item0[p], item0[p + 1], nnz[p], npan[p], real[p]};
    std::vector<int32_t> it, sg;
    xcd_lists(mem, it, sg);
    std::memcpy(items, it.data(), it.size() * 4);
    std::memcpy(seg, sg.data(), sg.size() * 4);
}
int hc_sign1(uint64_t c, int n, int nt) { return sign1(c, n, nt); }

// Full Hamiltonian through the mirror.  Returns nnz (or -nnz when cap is too
// small).  ql_major = 0: the build kernel's order (each open pair hop of the
// row word with its translated images summed by pair_orbit_sum, then the
// tunnelling hops per translation, then the diagonal); 1: the reference's
// loop order (translation, then pair and shift, through pair_term), which
// checks that regrouping.
static int64_t hamil_mirror(const blqh_sector* s, const blqh_params* p, int64_t* rowptr, int64_t* col, double* val,
                            int64_t cap, int ql_major) {
    HostBasis B;
    build(*s, B);
    HostTables T = host_tables(*s, *p);
    RowCtx c{};
    c.g = B.g;
    c.spec = B.spec;
    c.M = B.M;
    c.K = s->K;
    set_row_species(c, *s);
    c.period = B.period.data();
    c.lut = B.lut.data();
    c.cnt = B.cnt.data();
    c.C = T.C.data();
    c.FT = T.FT.data();
    c.tw = T.tw.data();
    const int nphi = s->nphi, nel = s->nel, Tn = 2 * nphi - 1;
    const int npairs = nel * (nel - 1) / 2;
    const int64_t n = (int64_t)B.id.size();
    std::vector<int64_t> rp(n + 1, 0), cl;
    std::vector<double> vl;
    auto occ_of = [&](uint64_t L, int* occ) {
        int k = 0;
        for (int b = 0; b < c.g.S * nphi; ++b)
            if ((L >> b) & 1ULL) occ[k++] = b;
    };
    for (int64_t i = 0; i < n; ++i) {
        std::map<int, std::pair<double, double>> acc;
        const uint64_t L0 = B.id[i];
        const int Dl = c.K < 0 ? 1 : B.period[i];
        uint64_t Ls[64];
        int sgl[64];
        for (int ql = 0; ql < Dl; ++ql) Ls[ql] = translate(c.g, L0, ql, sgl[ql]);
        auto add = [&](int j, double re, double im) {
            acc[j].first += re;
            acc[j].second += im;
        };
        if (ql_major) {
            for (int ql = 0; ql < Dl; ++ql) {
                int occ[64];
                occ_of(Ls[ql], occ);
                for (int cnd = 0; cnd < npairs * Tn; ++cnd) {
                    const int pr = cnd / Tn, t = cnd - pr * Tn - (nphi - 1);
                    const unsigned short pw = pair_decode(nel, pr);
                    int j;
                    double re, im;
                    if (pair_term(c, Ls[ql], sgl[ql], Dl, ql, occ[pw & 0xff], occ[pw >> 8], t, j, re, im)) add(j, re, im);
                }
                for (const HopTerm& h : T.hops) {
                    int j;
                    double re, im;
                    if (hop_term(c, Ls[ql], sgl[ql], Dl, ql, h, j, re, im)) add(j, re, im);
                }
            }
        } else {
            int occ[64];
            occ_of(L0, occ);
            for (int cnd = 0; cnd < npairs * Tn; ++cnd) {
                const int pr = cnd / Tn, t = cnd - pr * Tn - (nphi - 1);
                const unsigned short pw = pair_decode(nel, pr);
                int j;
                double re, im;
                if (pair_orbit_sum(c, Ls, sgl, Dl, occ[pw & 0xff], occ[pw >> 8], t, j, re, im)) add(j, re, im);
            }
            for (int ql = 0; ql < Dl; ++ql)
                for (const HopTerm& h : T.hops) {
                    int j;
                    double re, im;
                    if (hop_term(c, Ls[ql], sgl[ql], Dl, ql, h, j, re, im)) add(j, re, im);
                }
        }
        acc[(int)i].first += diag_term(c.g, c.spinless != 0, L0, T.Ec * nel, p->d, p->delta_v, p->delta_z);
        for (auto& kv : acc)
            if (std::hypot(kv.second.first, kv.second.second) > 1e-10) {
                cl.push_back(kv.first);
                vl.push_back(kv.second.first);
                vl.push_back(kv.second.second);
            }
        rp[i + 1] = (int64_t)cl.size();
    }
    const int64_t nnz = (int64_t)cl.size();
    if (nnz > cap) return -nnz;
    std::memcpy(rowptr, rp.data(), (n + 1) * 8);
    std::memcpy(col, cl.data(), nnz * 8);
    std::memcpy(val, vl.data(), nnz * 16);
    return nnz;
}

int64_t hc_hamil(const blqh_sector* s, const blqh_params* p, int64_t* rowptr, int64_t* col, double* val, int64_t cap) {
    return hamil_mirror(s, p, rowptr, col, val, cap, 0);
}

int64_t hc_hamil_ql_major(const blqh_sector* s, const blqh_params* p, int64_t* rowptr, int64_t* col, double* val,
                          int64_t cap) {
    return hamil_mirror(s, p, rowptr, col, val, cap, 1);
}

// The SpMV formats' internal order (panel.hip internal_order) of a basis given
// by its ids: out[i] = internal position of row i, or -1 everywhere for the
// identity.
void hc_internal_order(const blqh_sector* s, const uint64_t* ids, int64_t n, int32_t* out) {
    blqh_basis B;
    B.sec = *s;
    B.geo = Geo{s->nspecies, s->nphi, s->nel};
    B.n = n;
    B.h_id.assign(ids, ids + n);
    const std::vector<int32_t>& o = internal_order(&B);
    for (int64_t i = 0; i < n; ++i) out[i] = o.empty() ? -1 : o[i];
}

// 1 when the replayed hash-map iteration order (basis.hip) equals the order
// of a real std::unordered_map filled the same way; *n = words.
int hc_hash_order_check(const blqh_sector* s, int64_t* n) {
    const std::vector<uint64_t> a = hash_iteration_order(*s);
    const std::vector<uint64_t> b = hash_iteration_order_map(*s);
    *n = (int64_t)a.size();
    return a == b ? 1 : 0;
}

// The level-by-level order (the device's algorithm) against a real map.
int hc_hash_order_levels_check(const blqh_sector* s, int64_t* n) {
    const std::vector<uint64_t> a = level_order(*s);
    const std::vector<uint64_t> b = hash_iteration_order_map(*s);
    *n = (int64_t)a.size();
    return a == b ? 1 : 0;
}

// The build's reciprocal modulo (rowterms.hpp mod_by_rcp) for s = 0 .. n-1
// with a given reciprocal r of d.
void hc_mod_by_rcp(int n, int d, float r, int32_t* out) {
    for (int s = 0; s < n; ++s) out[s] = mod_by_rcp(s, d, r);
}

// the batched solver's low Ritz vectors (ritz.hpp)
void hc_tridiag_vector_k(int m, const double* a, const double* b, int k, double* x) {
    tridiag_vector_k(m, a, b, k, x);
}

void hc_excited_mix(int m, const double* a, const double* b, const double* c0, double* out) {
    excited_mix(m, a, b, c0, out);
}

}  // extern "C"
