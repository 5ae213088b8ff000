// engine.hpp -- internal types shared by the basis, Hamiltonian and Lanczos
// stages of the blqh library.
#pragma once
#include <hip/hip_runtime.h>
#include <atomic>
#include <cstdint>
#include <mutex>
#include <utility>
#include <vector>

#include "bits.hpp"
#include "common.hpp"
#include "panel.hpp"
#include "rowterms.hpp"

struct blqh_basis {
    blqh_sector sec{};
    blqh::Geo geo{};
    blqh::RankSpec spec{};           // the sector (with the spin-up constraint)
    blqh::DevBuf<uint32_t> cnt;      // rank table of spec
    int64_t M = 0;                   // sector words
    int64_t n = 0;                   // basis states (orbit representatives)
    blqh::DevBuf<uint64_t> id;       // representatives, ascending
    blqh::DevBuf<int16_t> period;    // orbit period (basis_C)
    blqh::DevBuf<int32_t> lut;       // sector word rank -> (row << 5) | shift q of its orbit member, -1 = none
    std::vector<uint64_t> h_id;
    std::vector<int16_t> h_period;
    double seconds_host = 0, seconds_device = 0;
    mutable std::atomic<int32_t> max_row_nnz{0};  // largest row kept by a build with the pair terms (0: none yet)
    // internal order of the SpMV formats (internal_order), computed once
    mutable std::mutex ord_mu;
    mutable bool ord_done = false;
    mutable std::vector<int32_t> h_order;  // h_order[i] = internal position of row i (empty: identity)
};

struct blqh_hamil {
    const blqh_basis* basis = nullptr;
    blqh_params par{};
    int64_t n = 0, nnz = 0;
    blqh::DevBuf<int64_t> rowptr;
    blqh::DevBuf<int32_t> col;
    blqh::DevBuf<double2> val;
    // Internal row / column order of every SpMV format (panels, tiles and the
    // CSR kernel): the CSR again with rows and columns permuted (internal_order).
    // Vectors cross into and out of it at the solver's boundaries.
    struct Order {
        bool on = false;
        blqh::DevBuf<int32_t> perm;      // perm[i]: internal position of reference row i
        blqh::DevBuf<int32_t> iperm;     // iperm[r]: reference row at internal position r
        std::vector<int32_t> h_iperm;
        blqh::DevBuf<int64_t> rowptr;    // the CSR in internal order, columns ascending
        blqh::DevBuf<int32_t> col;
        blqh::DevBuf<double2> val;
    } ord;
    blqh::Panels pan;                // panel-major copy for the SpMV hot loop
    blqh::Tiles til;                 // column-tile copy (SpMV kind 4), built on demand
    int spmv_kind = -1;              // autotuned SpMV: 0 = row-major CSR, 1 = panel-major (-1 = not tuned)
    int spmv_lpr = 0;                // lanes per row of the chosen kernel
    int spmv_threads = 0;            // workgroup size of the panel kernel
    int spmv_wgs_per_cu = 0;         // persistent workgroups per CU (panel kernel)
    int spmv_rows = 0;               // rows per lane group in flight (panel kernel)
    double seconds_build = 0;
    int terms = 3;                   // 1: Coulomb + diagonal, 2: interlayer tunnelling (unit Delta_SAS), 3: both
};

namespace blqh {

// Translation geometry of a sector: the bilayer trees translate by one
// orbital (nphi units), the monolayer by nphi / gcd(nel, nphi) orbitals
// (gcd(nel, nphi) units, Monolayer/basis.cpp:46-58).
inline Geo make_geo(const blqh_sector& s) {
    Geo g{s.nspecies, s.nphi, s.nel};
    if (s.nspecies == 1) {
        int c = 1;
        for (int i = 1; i <= s.nel; ++i)
            if (s.nel % i == 0 && s.nphi % i == 0) c = i;
        g.step = s.nphi / c;
    }
    return g;
}

// Species-dependent switches of the row terms.
inline void set_row_species(RowCtx& c, const blqh_sector& s) {
    c.spinless = s.nspecies == 2;
    c.twist_sel = s.nspecies == 2 ? 1 : 2;
    c.mono_odd = s.nspecies == 1 && (s.nphi & 1);
}

RankSpec make_rank_spec(const blqh_sector& s, bool with_spin_constraint);
std::vector<uint32_t> make_rank_table(const RankSpec& spec, int64_t* total);
void exclusive_scan(const int32_t* in, int64_t* out, int64_t n, hipStream_t st);  // out has n+1 entries
std::vector<uint64_t> hash_iteration_order(const blqh_sector& s);
std::vector<uint64_t> hash_iteration_order_map(const blqh_sector& s);  // from a real std::unordered_map
// rehash schedule of the map over n inserts: (first insert after the rehash, buckets)
std::vector<std::pair<int64_t, uint64_t>> hash_levels(int64_t n);
RankSpec reversed_rank_spec(const blqh_sector& s, const RankSpec& smap);  // bit-reversed words

// Event timing of the fused Lanczos SpMV (blqh_profile_*).
struct Profile {
    bool on = false;
    int64_t launches = 0;
    double ms = 0, bytes = 0;
    std::vector<hipEvent_t> ev;  // start/stop pairs, reused across solves
};
Profile& profile();

// Host-side tables of one Hamiltonian (Coulomb table, background energy,
// translation phases, twist phases, tunnelling hop list).
struct HostTables {
    std::vector<double> C;
    std::vector<double2> FT, tw;
    std::vector<HopTerm> hops;
    double Ec = 0;
};
HostTables host_tables(const blqh_sector& s, const blqh_params& p);
void build_panels(blqh_hamil* H);
// The SpMV formats of H (internal-order CSR, panels): built with a full
// Hamiltonian or a merge, on first use (or blqh_hamil_prepare) for the parts
// of a merge (terms 1 and 2), which often only feed blqh_hamil_combine or a
// shared-scale batch and never need them.
void ensure_panels(blqh_hamil* H);
// Column-tile layout for SpMV kind 4 on `G` workgroups; returns false (and
// builds nothing) when the matrix does not fit it.
bool build_tiles(blqh_hamil* H, int G);
// Converts the panel runs to (jagged = true) or from step-major order in place
// (the panel buffers keep their addresses); no-op when already there.  The
// step-major order of set_panel_layout is the greedy one (steps that share x
// cache lines across rows; it sizes its LDS from the runs, allocates and
// synchronises).  With greedy = false and a scratch area of at least
// panel_layout_scratch_bytes the plain step-major conversion is
// stream-ordered and allocates nothing; without a scratch area it allocates a
// temporary and synchronises the library stream.
// rows > 1: step-major with that many rows per lane (plain order), for SpMV kind 5.
void set_panel_layout(blqh_hamil* H, bool jagged, void* scratch = nullptr, size_t scratch_bytes = 0, int rows = 1);
bool greedy_step_order();  // single-matrix jagged panels in the greedy step order (BLQH_JAGGED_ORDER=greedy)
size_t panel_layout_scratch_bytes(const blqh_hamil* H);
// The same for many matrices in one launch (those already in the layout are skipped).
void set_panel_layouts(blqh_hamil* const* hs, int count, bool jagged, void* scratch, size_t scratch_bytes,
                       bool greedy, int rows = 1);
size_t panel_layout_scratch_bytes(blqh_hamil* const* hs, int count);

// The CSR that the SpMV formats are built from and the CSR kernel reads:
// the internal order's when it is on, else the Hamiltonian's own.
struct CsrView {
    const int64_t* rowptr;
    const int32_t* col;
    const double2* val;
};
inline CsrView spmv_csr(const blqh_hamil* H) {
    if (H->ord.on) return CsrView{H->ord.rowptr.get(), H->ord.col.get(), H->ord.val.get()};
    return CsrView{H->rowptr.get(), H->col.get(), H->val.get()};
}
// Internal order: a permutation of the basis rows (empty = identity) and the
// moves of vectors between the reference order and the internal one.
const std::vector<int32_t>& internal_order(const blqh_basis* B);
void build_order(blqh_hamil* H);
void order_to_internal(const blqh_hamil* H, const double2* src, double2* dst, hipStream_t st);
void order_to_reference(const blqh_hamil* H, const double2* src, double2* dst, hipStream_t st);

// Row context with the basis' device pointers (tables left null).
inline RowCtx make_row_ctx(const blqh_basis* B) {
    RowCtx c{};
    c.g = B->geo;
    c.spec = B->spec;
    c.M = B->M;
    c.K = B->sec.K;
    set_row_species(c, B->sec);
    c.period = B->period.get();
    c.lut = B->lut.get();
    c.cnt = B->cnt.get();
    return c;
}

}  // namespace blqh
