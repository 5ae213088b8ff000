"""The GPU path's per-element arithmetic, run on the CPU through the host
mirror (tests/hostcheck), compared with the oracle: basis bit-exact, CSR
structure exact, values to 1e-13 relative."""
import os

import numpy as np
import pytest

import blqh_amd
from tests import hostmirror

pytestmark = pytest.mark.skipif(not os.path.exists(hostmirror.PATH), reason="host mirror not built")

CASES = [
    ((2, 8, 8, 4, -1, 4, 4), dict(d=0.0)),
    ((2, 8, 8, 4, -1, 4, 4), dict(d=3.0)),
    ((2, 8, 8, -1, -1, 3, 2), dict(d=1.0, dsas=0.1, dv=0.05)),
    ((2, 10, 10, 5, -1, 1, 7), dict(d=0.5)),
    ((2, 8, 8, -1, -1, 0, -1), dict(d=1.0, dsas=0.2)),
    ((2, 9, 6, -1, -1, 2, 1), dict(d=0.7, dsas=0.1)),
    ((2, 8, 6, 2, -1, 1, 4), dict(d=1.2, dv=0.1)),
    ((4, 6, 6, 3, 3, 0, 1), dict(d=1.0, dsas=0.05, dv=0.02, dz=0.01)),
    ((4, 6, 6, -1, -1, 1, 2), dict(d=1.5, dsas=0.05, dv=0.02, dz=0.01)),
    ((4, 4, 6, 3, -1, 2, 3), dict(d=0.8, dsas=0.3, dz=0.05)),
]


def _params(nphi, kw):
    import oracle as O
    op = O.params(nphi, **kw)
    return op, blqh_amd.Params(op.lx, op.ly, op.nLL, op.d, op.dsas, op.dv, op.dz, 0, 0, 0)


@pytest.mark.parametrize("case,kw", CASES)
def test_mirror_matches_oracle(oracle_mod, case, kw):
    O = oracle_mod
    B = O.Basis(O.sector(*case))
    sec = blqh_amd.Sector(*case)
    ids, per = hostmirror.basis(sec)
    np.testing.assert_array_equal(ids, B.id)
    np.testing.assert_array_equal(per, B.period)
    op, par = _params(case[1], kw)
    H = O.Hamiltonian(B, op)
    rp, col, val = hostmirror.hamil(sec, par, len(B), 2 * H.nnz + 64)
    np.testing.assert_array_equal(rp, H.rowptr)
    np.testing.assert_array_equal(col, H.col)
    np.testing.assert_allclose(val, H.val, rtol=0, atol=1e-13 * np.abs(H.val).max())


@pytest.mark.parametrize("case", [(2, 8, 8, -1, -1, 3, 0), (2, 8, 8, 4, -1, -1, -1), (4, 6, 6, 3, 3, 5, 0),
                                  (4, 6, 6, 3, -1, 2, 0), (2, 7, 5, -1, -1, 6, 0)])
def test_rank_unrank_enumerates_sector(oracle_mod, case):
    """unrank(0..M-1) lists the sector ascending; rank inverts it."""
    O = oracle_mod
    sec = blqh_amd.Sector(*case)
    words = hostmirror.sector_words(sec, with_spin=True)
    ref = np.sort(O.sector_configs(O.sector(*case)))
    if case[0] == 4 and case[4] >= 0:
        nphi = case[1]
        m = (1 << nphi) - 1
        up = np.array([bin(int(w) & m).count("1") + bin((int(w) >> (2 * nphi)) & m).count("1") for w in ref])
        ref = ref[up == case[4]]
    np.testing.assert_array_equal(words, ref)
