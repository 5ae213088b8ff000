// ref_basis_driver.cpp -- C entry points around the reference's own basis
// class, compiled together with the reference's basis.cpp (spinless or
// spinful variant, selected by the include path) into oracle/_ref/.
//
// TEST INFRASTRUCTURE ONLY: used by tests/ to pin the oracle's basis
// (representative choice, index order, orbit periods, translation and hop
// signs) against the reference's code built from its own sources.
#include <cstdint>
#include "basis.h"

extern "C" {

// Build a sector with the reference class and copy id[] / basis_C[] out.
// Returns nbasis, or -nbasis if cap is too small.
int64_t ref_basis(int nphi, int nel, int nel_up, int nel_su, int J, int K,
                  uint64_t* id, int16_t* period, int64_t cap) {
    basis b;
#ifdef REF_SPINFUL
    if (nel_up >= 0) b.init(nphi, nel, nel_up, nel_su, J, K);
    else b.init(nphi, nel, J, K);
#else
    (void)nel_su;
    b.init(nphi, nel, nel_up, J, K);
#endif
    int64_t n = (int64_t)b.id.size();
    if (n > cap) return -n;
    for (int64_t i = 0; i < n; ++i) {
        id[i] = b.id[i];
        period[i] = b.basis_C[i];
    }
    return n;
}

// Translation primitives on a sector-initialised object (popcount table).
static basis* g_b = nullptr;
int ref_prepare(int nphi, int nel, int nel_up, int nel_su, int J, int K) {
    delete g_b;
    g_b = new basis();
#ifdef REF_SPINFUL
    if (nel_up >= 0) g_b->init(nphi, nel, nel_up, nel_su, J, K);
    else g_b->init(nphi, nel, J, K);
#else
    (void)nel_su;
    g_b->init(nphi, nel, nel_up, J, K);
#endif
    return (int)g_b->id.size();
}
uint64_t ref_translate(uint64_t c, int k, int* sign) { return g_b->translate(c, k, *sign); }
uint64_t ref_inv_translate(uint64_t c, int k, int* sign) { return g_b->inv_translate(c, k, *sign); }
int ref_sign2(uint64_t c, int n, int m, int nt, int mt) { return g_b->get_sign(c, n, m, nt, mt, 0); }
int ref_sign1(uint64_t c, int n, int nt) { return g_b->get_sign(c, n, nt); }

}
