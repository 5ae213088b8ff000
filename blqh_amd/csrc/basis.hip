// basis.hip -- sector basis on the GPU.
//
// Pipeline (reference: basis::init, BLQH_without_spin/basis.cpp:61-120 and
// BLQH_with_spin/basis.cpp:74-141):
//  1. The reference inserts the sector words into a
//     std::unordered_map<unsigned long, long> in depth-first order and keeps,
//     for every translation orbit, the member that comes last in the map's
//     iteration order (it erases a word while iterating whenever another
//     orbit member is still present), so that order is the only input the
//     representative choice needs.  The host derives the container's rehash
//     schedule; the device generates the words in insertion order and
//     computes every word's list position level by level (map_positions).
//  2. device: combinatorial rank of every word -> its position in that order.
//  3. device: survivor test per word (no orbit member later in the order),
//     spin-up filter, compaction in rank order = ascending word order (the
//     reference's sort of id[]).
//  4. device: orbit periods (basis_C) and the lookup table used by the
//     Hamiltonian build: for every sector word w, the representative row j
//     and the shift q with inv_translate(w, q) == id[j].
#include <algorithm>
#include <chrono>
#include <unordered_map>

#include "engine.hpp"

namespace blqh {

RankSpec make_rank_spec(const blqh_sector& s, bool with_spin_constraint) {
    RankSpec r{};
    r.nbits = s.nspecies * s.nphi;
    r.nphi = s.nphi;
    r.nel = s.nel;
    r.R = s.J >= 0 ? s.nphi : 1;
    r.Jt = s.J >= 0 ? s.J : 0;
    r.maskA = s.nel_up >= 0 ? lowbits(r.nbits / 2) : 0;
    r.nA = s.nel_up >= 0 ? s.nel_up : 0;
    if (with_spin_constraint && s.nspecies == 4 && s.nel_su >= 0) {
        r.maskB = lowbits(s.nphi) | (lowbits(s.nphi) << (2 * s.nphi));
        r.nB = s.nel_su;
    } else {
        r.maskB = 0;
        r.nB = 0;
    }
    r.dK = s.nel + 1;
    r.dR = r.R;
    r.dA = r.maskA ? r.nA + 1 : 1;
    r.dB = r.maskB ? r.nB + 1 : 1;
    r.sA = r.dB;
    r.sR = (int64_t)r.dA * r.sA;
    r.sK = (int64_t)r.dR * r.sR;
    r.sP = (int64_t)r.dK * r.sK;
    return r;
}

// cnt[p][k][r][a][b]: subsets of [0,p) with k elements, residual r, a in A,
// b in B.  Entries are saturated to 32 bits; the rank/unrank walks only read
// entries bounded by the sector size, which is checked to fit.
std::vector<uint32_t> make_rank_table(const RankSpec& s, int64_t* total) {
    const int64_t sz = (int64_t)(s.nbits + 1) * s.sP;
    BLQH_REQUIRE(sz < (int64_t)INT32_MAX, BLQH_ERANGE, "rank table too large");  // rank_word indexes in 32 bits
    std::vector<int64_t> c(sz, 0);
    auto sat_add = [](int64_t x, int64_t y) { int64_t z = x + y; return z < x ? INT64_MAX : z; };
    c[rs_index(s, 0, 0, 0, 0, 0)] = 1;
    for (int p = 0; p < s.nbits; ++p) {
        const int inA = (s.maskA >> p) & 1, inB = (s.maskB >> p) & 1;
        for (int k = 0; k < s.dK; ++k)
            for (int r = 0; r < s.dR; ++r)
                for (int a = 0; a < s.dA; ++a)
                    for (int b = 0; b < s.dB; ++b) {
                        int64_t v = c[rs_index(s, p, k, r, a, b)];
                        if (k > 0 && a - inA >= 0 && b - inB >= 0) {
                            int rr = (r - p % s.nphi) % s.R;
                            if (rr < 0) rr += s.R;
                            v = sat_add(v, c[rs_index(s, p, k - 1, rr, a - inA, b - inB)]);
                        }
                        c[rs_index(s, p + 1, k, r, a, b)] = v;
                    }
    }
    *total = c[rs_index(s, s.nbits, s.nel, s.Jt, s.nA, s.nB)];
    std::vector<uint32_t> out(sz);
    for (int64_t i = 0; i < sz; ++i) out[i] = c[i] > 0xFFFFFFFFLL ? 0xFFFFFFFFu : (uint32_t)c[i];
    return out;
}

}  // namespace blqh

namespace {

using namespace blqh;

constexpr int kBlock = 256;

int grid_for(int64_t n, int block = kBlock) {
    int64_t g = ceil_div(n, block);
    return (int)std::min<int64_t>(std::max<int64_t>(g, 1), 1 << 20);
}

// ---- scan ------------------------------------------------------------------
constexpr int kScanBlock = 1024;

__global__ void k_block_sums(const int32_t* __restrict__ in, int64_t n, int64_t* __restrict__ sums) {
    __shared__ int64_t sh[kScanBlock / 64];
    const int64_t i = (int64_t)blockIdx.x * kScanBlock + threadIdx.x;
    int64_t v = i < n ? in[i] : 0;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t t = 0;
        for (int w = 0; w < kScanBlock / 64; ++w) t += sh[w];
        sums[blockIdx.x] = t;
    }
}

// single-block exclusive scan of the block sums (in place), total to *tot
__global__ void k_scan_sums(int64_t* __restrict__ sums, int64_t nb, int64_t* __restrict__ tot) {
    __shared__ int64_t carry;
    __shared__ int64_t sh[kScanBlock];
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int64_t base = 0; base < nb; base += kScanBlock) {
        const int64_t i = base + threadIdx.x;
        int64_t v = i < nb ? sums[i] : 0;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < kScanBlock; o <<= 1) {
            int64_t t = threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
            __syncthreads();
            sh[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nb) sums[i] = carry + sh[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == kScanBlock - 1) carry += sh[threadIdx.x];
        __syncthreads();
    }
    if (threadIdx.x == 0) *tot = carry;
}

__global__ void k_block_scan(const int32_t* __restrict__ in, int64_t n, const int64_t* __restrict__ offs,
                             int64_t* __restrict__ out) {
    __shared__ int64_t sh[kScanBlock];
    const int64_t i = (int64_t)blockIdx.x * kScanBlock + threadIdx.x;
    const int64_t v = i < n ? in[i] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < kScanBlock; o <<= 1) {
        int64_t t = threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    if (i < n) out[i] = offs[blockIdx.x] + sh[threadIdx.x] - v;
}

__global__ void k_copy_total(const int64_t* __restrict__ tot, int64_t* __restrict__ dst) { *dst = *tot; }

// ---- basis kernels -----------------------------------------------------------

// flags[r] = 1 when word r is the orbit member latest in the hash order (or
// when no reduction is requested) and passes the spin-up filter.
__global__ void k_survivor(int64_t M, RankSpec s, const uint32_t* __restrict__ cnt,
                           const int32_t* __restrict__ posOf, Geo g, int reduce, uint64_t spin_mask, int nel_su,
                           int32_t* __restrict__ flags) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < M; r += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t c = unrank_word(s, cnt, r);
        bool keep = !reduce || is_survivor(g, s, cnt, posOf, c, posOf[r]);
        if (keep && nel_su >= 0 && popc(c & spin_mask) != nel_su) keep = false;
        flags[r] = keep ? 1 : 0;
    }
}

__global__ void k_compact(int64_t M, RankSpec s, const uint32_t* __restrict__ cnt, const int32_t* __restrict__ flags,
                          const int64_t* __restrict__ offs, uint64_t* __restrict__ id) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < M; r += (int64_t)gridDim.x * blockDim.x)
        if (flags[r]) id[offs[r]] = unrank_word(s, cnt, r);
}

// basis_C: smallest n in [1, order) with inv_translate(id, n) == id, else the
// order of the translation group (nphi in the bilayer trees, gcd(nel, nphi)
// in the monolayer).
__global__ void k_period(int64_t n, const uint64_t* __restrict__ id, Geo g, int16_t* __restrict__ period) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t c = id[j];
        const int ord = g.order();
        int per = ord;
        for (int k = 1; k < ord; ++k) {
            int sg;
            if (inv_translate(g, c, k, sg) == c) { per = k; break; }
        }
        period[j] = (int16_t)per;
    }
}

__global__ void k_fill_i32(int32_t* __restrict__ p, int64_t n, int32_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// lut[rank(translate(id[j], q))] = j << 5 | q for the orbit members in the sector
__global__ void k_lut(int64_t n, const uint64_t* __restrict__ id, const int16_t* __restrict__ period, Geo g,
                      RankSpec s, const uint32_t* __restrict__ cnt, int no_reduction, int32_t* __restrict__ lut) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t c = id[j];
        const int D = no_reduction ? 1 : period[j];
        for (int q = 0; q < D; ++q) {
            int sg;
            const uint64_t y = translate(g, c, q, sg);
            if (!in_sector(s, y)) continue;
            lut[rank_word(s, cnt, y)] = lut_pack(j, q);
        }
    }
}

// ---- the map's iteration order, level by level -----------------------------
//
// libstdc++'s unordered_map keeps its nodes in one singly linked list in which
// the nodes of a bucket are contiguous.  An insert puts the node at the front
// of its bucket's run, or at the front of the list when the bucket is empty;
// a rehash walks the old list from the front and applies the same rule with
// the new bucket count.  Give every node a time: its position in the list
// before the last rehash (0, 1, ...), or its insertion index for a node
// inserted after it (larger than every old position).  Then the list orders
// the buckets by decreasing earliest time of their nodes and the nodes of a
// bucket by decreasing time, so after each level of the rehash schedule
//   pos(x) = (nodes of the buckets whose earliest time is later than that of
//             x's bucket) + (nodes of x's bucket with a later time than x),
// which the three kernels below compute in parallel.

// key[i] = i-th word of the depth-first insertion order.  That order is
// lexicographic in the ascending orbital list, i.e. descending in the
// bit-reversed word, so it is the reversed sector (srev) unranked from the top.
__global__ void k_dfs_keys(int64_t M, RankSpec srev, const uint32_t* __restrict__ cnt, int nbits,
                           uint64_t* __restrict__ key) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < M; i += (int64_t)gridDim.x * blockDim.x)
        key[i] = __builtin_bitreverse64(unrank_word(srev, cnt, M - 1 - i)) >> (64 - nbits);
}

// Level step 1 (nodes [0, en), of which [st, en) are inserted after the
// rehash to m buckets): time of the new nodes, bucket, bucket size and
// earliest time.
__global__ void k_hl_bucket(int64_t st, int64_t en, uint64_t m, const uint64_t* __restrict__ key,
                            uint32_t* __restrict__ tau, uint32_t* __restrict__ bucket, int32_t* __restrict__ gcount,
                            uint32_t* __restrict__ gmin) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < en; i += (int64_t)gridDim.x * blockDim.x) {
        if (i >= st) tau[i] = (uint32_t)i;
        const uint32_t b = (uint32_t)(key[i] % m);
        bucket[i] = b;
        atomicAdd(&gcount[b], 1);
        atomicMin(&gmin[b], tau[i]);
    }
}

// Step 2 (goff = exclusive scan of the bucket sizes): each bucket's nodes,
// in any order, and the size of every bucket at the slot of its earliest
// time counted from the latest time down (first_sz[en - 1 - time]).
// gcount returns to zero.
__global__ void k_hl_members(int64_t en, const uint32_t* __restrict__ tau, const uint32_t* __restrict__ bucket,
                             const int64_t* __restrict__ goff, int32_t* __restrict__ gcount,
                             const uint32_t* __restrict__ gmin, uint32_t* __restrict__ members,
                             int32_t* __restrict__ first_sz) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < en; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t b = bucket[i];
        members[goff[b] + atomicSub(&gcount[b], 1) - 1] = (uint32_t)i;
        first_sz[en - 1 - tau[i]] = tau[i] == gmin[b] ? (int32_t)(goff[b + 1] - goff[b]) : 0;
    }
}

// Step 3 (later = exclusive scan of first_sz: later[en - 1 - t] counts the
// nodes of the buckets whose earliest time exceeds t): list positions.
__global__ void k_hl_pos(int64_t en, const uint32_t* __restrict__ tau, const uint32_t* __restrict__ bucket,
                         const int64_t* __restrict__ goff, const uint32_t* __restrict__ gmin,
                         const uint32_t* __restrict__ members, const int64_t* __restrict__ later,
                         uint32_t* __restrict__ tau_out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < en; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t b = bucket[i], t = tau[i];
        int64_t p = later[en - 1 - gmin[b]];
        for (int64_t k = goff[b]; k < goff[b + 1]; ++k) p += tau[members[k]] > t ? 1 : 0;
        tau_out[i] = (uint32_t)p;
    }
}

// posOf[rank(key[i])] = final list position of node i
__global__ void k_scatter_tau(const uint64_t* __restrict__ key, const uint32_t* __restrict__ tau, int64_t M,
                              RankSpec s, const uint32_t* __restrict__ cnt, int32_t* __restrict__ posOf) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < M; i += (int64_t)gridDim.x * blockDim.x)
        posOf[rank_word(s, cnt, key[i])] = (int32_t)tau[i];
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, int64_t n, uint32_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ---- host helpers ------------------------------------------------------------

// The reference's depth-first insertion order (generate / generate_all_density /
// generate_fixed_density): electron k scans its allowed orbital window upward,
// the first nel_up electrons in the upper layer, the rest in the lower one.
template <class F>
void dfs_words(const blqh_sector& s, F&& emit) {
    const int nbits = s.nspecies * s.nphi, W = nbits / 2, ne = s.nel;
    if (ne == 0) { if (s.J <= 0) emit(0ULL); return; }
    std::vector<int> lo(ne), hi(ne), pos(ne);
    for (int k = 0; k < ne; ++k) {
        if (s.nel_up >= 0 && k < s.nel_up) { lo[k] = 0; hi[k] = W - (s.nel_up - k - 1); }
        else if (s.nel_up >= 0) { lo[k] = W; hi[k] = nbits - (ne - k - 1); }
        else { lo[k] = 0; hi[k] = nbits - (ne - k - 1); }
    }
    std::vector<uint64_t> word(ne + 1, 0);
    std::vector<int> js(ne + 1, 0);
    int k = 0;
    pos[0] = lo[0];
    while (k >= 0) {
        if (pos[k] >= hi[k]) { if (--k >= 0) ++pos[k]; continue; }
        const uint64_t c = word[k] | (1ULL << pos[k]);
        const int j = js[k] + pos[k] % s.nphi;
        if (k + 1 == ne) {
            if (s.J < 0 || j % s.nphi == s.J) emit(c);
            ++pos[k];
        } else {
            word[k + 1] = c;
            js[k + 1] = j;
            pos[k + 1] = std::max(pos[k] + 1, lo[k + 1]);
            ++k;
        }
    }
}

void validate_sector(const blqh_sector& s) {
    BLQH_REQUIRE(s.nspecies == 1 || s.nspecies == 2 || s.nspecies == 4, BLQH_EINVAL, "nspecies must be 1, 2 or 4");
    BLQH_REQUIRE(s.nspecies != 1 || s.nel_up < 0, BLQH_EINVAL, "a monolayer sector has no layer occupation (nel_up = -1)");
    BLQH_REQUIRE(s.nphi >= 2 && s.nphi <= 31 && s.nspecies * s.nphi <= 64, BLQH_EINVAL,
                 "nphi must satisfy 2 <= nphi <= 31 and nspecies*nphi <= 64");
    BLQH_REQUIRE(s.nel >= 1 && s.nel <= s.nspecies * s.nphi, BLQH_EINVAL, "nel out of range");
    BLQH_REQUIRE(s.nel_up <= s.nel, BLQH_EINVAL, "nel_up must not exceed nel");
    BLQH_REQUIRE(s.nel_up < 0 || (s.nel_up <= s.nspecies * s.nphi / 2 && s.nel - s.nel_up <= s.nspecies * s.nphi / 2),
                 BLQH_EINVAL, "layer occupation exceeds the layer");
    BLQH_REQUIRE(s.J < s.nphi && s.K < s.nphi, BLQH_EINVAL, "J and K must be < nphi");
    BLQH_REQUIRE(s.nspecies != 1 || s.K < make_geo(s).order(), BLQH_EINVAL, "monolayer K must be < gcd(nel, nphi)");
    BLQH_REQUIRE(s.nel_su < 0 || s.nspecies == 4, BLQH_EINVAL, "nel_su needs nspecies = 4");
}

}  // namespace

namespace blqh {

// Iteration order of the reference's std::unordered_map<unsigned long, long>
// after inserting the map sector (no spin-up filter yet) in depth-first order,
// on the host.  The basis takes the order from the device computation
// (map_positions); this sequential replay is the CPU tests' check of it.
//
// The map is not built: its node list is replayed with index arrays.  The
// container keeps every node in one singly linked list; bucket b (key mod
// the bucket count) holds the node before its first node, or the list head
// marker.  A key whose bucket is occupied goes right after that node; a key
// of an empty bucket goes to the front of the list (and the bucket of the
// former first node now starts after it).  A rehash walks the old list and
// re-inserts each node the same way into the empty larger table.  When to
// rehash and to how many buckets is decided by the library's own policy
// object, called exactly as the container calls it, so the bucket counts are
// this libstdc++'s.  Every word is new (the enumeration has no repeats), so
// no lookup precedes an insert.  Same order as the real map (tested on every
// test sector, hash_iteration_order_map); about 4x faster at 37.6M words.
std::vector<uint64_t> hash_iteration_order(const blqh_sector& s) {
    std::vector<uint64_t> key;
    dfs_words(s, [&](uint64_t c) { key.push_back(c); });
    const int64_t n = (int64_t)key.size();
    BLQH_REQUIRE(n < ((int64_t)1 << 31) - 2, BLQH_ERANGE, "sector too large for the hash-order replay");
    constexpr int32_t kEmpty = -1, kHead = -2;  // bucket empty / bucket starts after the list head
    std::vector<int32_t> nxt(n, -1), bkt(1, kEmpty), nb;
    int32_t head = -1;
    uint64_t nbkt = 1;
    std::__detail::_Prime_rehash_policy policy;  // the container's policy, default max load factor 1
    auto next_of = [&](int32_t p) { return p == kHead ? head : nxt[p]; };
    auto set_next = [&](int32_t p, int32_t v) {
        if (p == kHead) head = v; else nxt[p] = v;
    };
    for (int64_t i = 0; i < n; ++i) {
        const std::pair<bool, std::size_t> rh = policy._M_need_rehash(nbkt, (std::size_t)i, 1);
        if (rh.first) {
            const uint64_t m = rh.second;
            nb.assign(m, kEmpty);
            int32_t p = head;
            head = -1;
            uint64_t bb = 0;
            while (p >= 0) {
                const int32_t q = nxt[p];
                const uint64_t b = key[p] % m;
                if (nb[b] == kEmpty) {
                    nxt[p] = head;
                    head = p;
                    nb[b] = kHead;
                    if (nxt[p] >= 0) nb[bb] = p;
                    bb = b;
                } else {
                    const int32_t prev = nb[b];
                    nxt[p] = next_of(prev);
                    set_next(prev, p);
                }
                p = q;
            }
            bkt.swap(nb);
            nbkt = m;
        }
        const int32_t node = (int32_t)i;
        const uint64_t b = key[i] % nbkt;
        if (bkt[b] != kEmpty) {
            const int32_t prev = bkt[b];
            nxt[node] = next_of(prev);
            set_next(prev, node);
        } else {
            nxt[node] = head;
            head = node;
            if (nxt[node] >= 0) bkt[key[nxt[node]] % nbkt] = node;
            bkt[b] = kHead;
        }
    }
    std::vector<uint64_t> order;
    order.reserve(n);
    for (int32_t p = head; p >= 0; p = nxt[p]) order.push_back(key[p]);
    BLQH_REQUIRE((int64_t)order.size() == n, BLQH_EINTERNAL, "hash-order replay lost a node");
    return order;
}

// The same order from a real std::unordered_map (the check of the replay).
std::vector<uint64_t> hash_iteration_order_map(const blqh_sector& s) {
    std::unordered_map<unsigned long, long> set;
    dfs_words(s, [&](uint64_t c) { set[(unsigned long)c] = (long)c; });
    std::vector<uint64_t> order;
    order.reserve(set.size());
    for (const auto& kv : set) order.push_back((uint64_t)kv.first);
    return order;
}

// The sector of the bit-reversed words (bit p -> nbits - 1 - p): the upper
// layer moves to the high half and orbital x of a block to nphi - 1 - x, so
// the momentum residual becomes nel (nphi - 1) - J.  smap has no spin-up
// constraint.
RankSpec reversed_rank_spec(const blqh_sector& s, const RankSpec& smap) {
    RankSpec srev = smap;
    if (smap.maskA) srev.maskA = lowbits(smap.nbits) & ~lowbits(smap.nbits / 2);
    if (smap.R > 1) srev.Jt = ((s.nel * (s.nphi - 1) - smap.Jt) % s.nphi + s.nphi) % s.nphi;
    return srev;
}

// Rehash schedule of the map while n words are inserted one by one: (index
// of the first insert after the rehash, bucket count), from the container's
// initial single bucket.
std::vector<std::pair<int64_t, uint64_t>> hash_levels(int64_t n) {
    std::__detail::_Prime_rehash_policy policy;  // default max load factor 1
    std::vector<std::pair<int64_t, uint64_t>> lv{{0, 1}};
    uint64_t nbkt = 1;
    for (int64_t i = 0; i < n; ++i) {
        const std::pair<bool, std::size_t> rh = policy._M_need_rehash(nbkt, (std::size_t)i, 1);
        if (rh.first) {
            nbkt = rh.second;
            lv.push_back({i, nbkt});
        }
    }
    return lv;
}

void exclusive_scan(const int32_t* in, int64_t* out, int64_t n, hipStream_t st) {
    const int64_t nb = std::max<int64_t>(1, ceil_div(n, kScanBlock));
    DevBuf<int64_t> sums(nb + 1);
    hipLaunchKernelGGL(k_block_sums, dim3((unsigned)nb), dim3(kScanBlock), 0, st, in, n, sums.get());
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanBlock), 0, st, sums.get(), nb, sums.get() + nb);
    hipLaunchKernelGGL(k_block_scan, dim3((unsigned)nb), dim3(kScanBlock), 0, st, in, n, sums.get(), out);
    hipLaunchKernelGGL(k_copy_total, dim3(1), dim3(1), 0, st, sums.get() + nb, out + n);
    BLQH_HIP(hipGetLastError());
    BLQH_HIP(hipStreamSynchronize(st));  // sums is freed on return
}

// posOf[rank(w)] = position of sector word w in the map's iteration order
// (see the level kernels above).  smap/cnt: the sector without the spin-up
// constraint, M its size.
void map_positions(const blqh_sector& s, const RankSpec& smap, const uint32_t* cnt, int64_t M, int32_t* posOf,
                   hipStream_t st) {
    BLQH_REQUIRE(M < ((int64_t)1 << 31) - 2, BLQH_ERANGE, "sector too large for the hash order");
    const RankSpec srev = reversed_rank_spec(s, smap);
    int64_t Mrev = 0;
    const std::vector<uint32_t> trev = make_rank_table(srev, &Mrev);
    BLQH_REQUIRE(Mrev == M, BLQH_EINTERNAL, "reversed sector size mismatch");
    const std::vector<std::pair<int64_t, uint64_t>> lv = hash_levels(M);
    uint64_t mmax = 1;
    for (const auto& l : lv) mmax = std::max(mmax, l.second);
    BLQH_REQUIRE(mmax < ((uint64_t)1 << 32), BLQH_ERANGE, "bucket count beyond 32 bits");
    DevBuf<uint32_t> dtrev(trev.size());
    BLQH_HIP(hipMemcpyAsync(dtrev.get(), trev.data(), trev.size() * 4, hipMemcpyHostToDevice, st));
    DevBuf<uint64_t> key(M);
    DevBuf<uint32_t> tau(M), tau2(M), bucket(M), members(M), gmin(mmax);
    DevBuf<int32_t> gcount(mmax), first_sz(M);
    DevBuf<int64_t> goff(mmax + 1), later(M + 1);
    hipLaunchKernelGGL(k_dfs_keys, dim3(grid_for(M)), dim3(kBlock), 0, st, M, srev, dtrev.get(), smap.nbits,
                       key.get());
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for((int64_t)mmax)), dim3(kBlock), 0, st, gcount.get(), (int64_t)mmax, 0);
    BLQH_HIP(hipGetLastError());
    uint32_t* cur = tau.get();
    uint32_t* nxt = tau2.get();
    for (size_t r = 0; r < lv.size(); ++r) {
        const int64_t st_r = lv[r].first, en = r + 1 < lv.size() ? lv[r + 1].first : M;
        const uint64_t m = lv[r].second;
        if (en == 0) continue;  // rehash before the first insert
        hipLaunchKernelGGL(k_fill_u32, dim3(grid_for((int64_t)m)), dim3(kBlock), 0, st, gmin.get(), (int64_t)m,
                           0xFFFFFFFFu);
        hipLaunchKernelGGL(k_hl_bucket, dim3(grid_for(en)), dim3(kBlock), 0, st, st_r, en, m, key.get(), cur,
                           bucket.get(), gcount.get(), gmin.get());
        BLQH_HIP(hipGetLastError());
        exclusive_scan(gcount.get(), goff.get(), (int64_t)m, st);
        hipLaunchKernelGGL(k_hl_members, dim3(grid_for(en)), dim3(kBlock), 0, st, en, cur, bucket.get(), goff.get(),
                           gcount.get(), gmin.get(), members.get(), first_sz.get());
        BLQH_HIP(hipGetLastError());
        exclusive_scan(first_sz.get(), later.get(), en, st);
        hipLaunchKernelGGL(k_hl_pos, dim3(grid_for(en)), dim3(kBlock), 0, st, en, cur, bucket.get(), goff.get(),
                           gmin.get(), members.get(), later.get(), nxt);
        BLQH_HIP(hipGetLastError());
        std::swap(cur, nxt);
    }
    hipLaunchKernelGGL(k_scatter_tau, dim3(grid_for(M)), dim3(kBlock), 0, st, key.get(), cur, M, smap, cnt, posOf);
    BLQH_HIP(hipGetLastError());
    BLQH_HIP(hipStreamSynchronize(st));
}

}  // namespace blqh

using namespace blqh;

// Build a basis; called through blqh_basis_create.
void blqh_basis_build_impl(const blqh_sector& s, blqh_basis* B) {
    validate_sector(s);
    hipStream_t st = stream();
    B->sec = s;
    B->geo = make_geo(s);
    const bool reduce = s.K >= 0;
    const bool spin_filter = s.nspecies == 4 && s.nel_su >= 0;

    // the words the reference inserts into its map: no spin-up constraint yet
    RankSpec smap = make_rank_spec(s, false);
    int64_t Mmap = 0;
    std::vector<uint32_t> tmap = make_rank_table(smap, &Mmap);
    BLQH_REQUIRE(Mmap > 0, BLQH_EINVAL, "empty sector");
    BLQH_REQUIRE(Mmap < (int64_t)INT32_MAX, BLQH_ERANGE, "sector too large (more than 2^31 words)");
    DevBuf<uint32_t> dmap(tmap.size());
    BLQH_HIP(hipMemcpyAsync(dmap.get(), tmap.data(), tmap.size() * 4, hipMemcpyHostToDevice, st));

    auto t0 = std::chrono::steady_clock::now();
    DevBuf<int32_t> posOf;
    if (reduce) {
        // every word's position in the map's iteration order after the
        // depth-first inserts (the rehash schedule on the host, the list
        // positions level by level on the device)
        posOf.alloc(Mmap);
        map_positions(s, smap, dmap.get(), Mmap, posOf.get(), st);
        B->seconds_host = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    auto t1 = std::chrono::steady_clock::now();
    DevBuf<int32_t> flags(Mmap);
    DevBuf<int64_t> offs(Mmap + 1);
    const uint64_t spin_mask = lowbits(s.nphi) | (lowbits(s.nphi) << (2 * s.nphi));
    hipLaunchKernelGGL(k_survivor, dim3(grid_for(Mmap)), dim3(kBlock), 0, st, Mmap, smap, dmap.get(), posOf.get(),
                       B->geo, reduce ? 1 : 0, spin_mask, spin_filter ? s.nel_su : -1, flags.get());
    BLQH_HIP(hipGetLastError());
    exclusive_scan(flags.get(), offs.get(), Mmap, st);
    int64_t n = 0;
    copy_sync(&n, offs.get() + Mmap, 8, hipMemcpyDeviceToHost);
    BLQH_REQUIRE(n > 0, BLQH_EINVAL, "empty basis");
    BLQH_REQUIRE(n < (1LL << 26), BLQH_ERANGE, "basis larger than 2^26 states is not supported");
    B->n = n;
    B->id.alloc(n);
    B->period.alloc(n);
    hipLaunchKernelGGL(k_compact, dim3(grid_for(Mmap)), dim3(kBlock), 0, st, Mmap, smap, dmap.get(), flags.get(),
                       offs.get(), B->id.get());
    hipLaunchKernelGGL(k_period, dim3(grid_for(n)), dim3(kBlock), 0, st, n, B->id.get(), B->geo, B->period.get());
    BLQH_HIP(hipGetLastError());

    // lookup table over the (spin-constrained) sector
    B->spec = make_rank_spec(s, true);
    std::vector<uint32_t> tsec = make_rank_table(B->spec, &B->M);
    B->cnt.alloc(tsec.size());
    BLQH_HIP(hipMemcpyAsync(B->cnt.get(), tsec.data(), tsec.size() * 4, hipMemcpyHostToDevice, st));
    B->lut.alloc(B->M);
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(B->M)), dim3(kBlock), 0, st, B->lut.get(), B->M, -1);
    hipLaunchKernelGGL(k_lut, dim3(grid_for(n)), dim3(kBlock), 0, st, n, B->id.get(), B->period.get(), B->geo,
                       B->spec, B->cnt.get(), reduce ? 0 : 1, B->lut.get());
    BLQH_HIP(hipGetLastError());
    B->h_id.resize(n);
    B->h_period.resize(n);
    BLQH_HIP(hipMemcpyAsync(B->h_id.data(), B->id.get(), n * 8, hipMemcpyDeviceToHost, st));
    BLQH_HIP(hipMemcpyAsync(B->h_period.data(), B->period.get(), n * 2, hipMemcpyDeviceToHost, st));
    BLQH_HIP(hipStreamSynchronize(st));
    B->seconds_device = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
}
