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


GROUPING = [
    ((2, 8, 8, 4, -1, 4, 4), dict(d=1.0)),
    ((2, 10, 10, -1, -1, 3, 7), dict(d=0.5, delta_sas=0.1, theta_x=0.7, theta_y=1.3)),
    ((2, 9, 6, -1, -1, 2, 1), dict(d=0.7, delta_sas=0.1, delta_v=0.05, theta_x=2.1)),
    ((4, 6, 6, 3, 3, 0, 1), dict(d=1.0, delta_sas=0.05, delta_v=0.02, delta_z=0.01, theta_b=0.4, theta_y=0.9)),
    ((4, 6, 6, -1, -1, 1, 2), dict(d=1.5, delta_sas=0.05, theta_x=1.1, theta_b=0.3)),
    ((1, 9, 3, -1, -1, 4, 2), dict()), ((1, 12, 4, -1, -1, 3, 3), dict()), ((1, 11, 4, -1, -1, 5, 0), dict()),
]


@pytest.mark.parametrize("case,kw", GROUPING, ids=lambda c: "_".join(map(str, c)) if isinstance(c, tuple) else None)
def test_orbit_grouped_terms_equal_reference_loop_order(case, kw):
    """The build kernel locates each open pair hop of the row word once and
    sums its images in the D_l translated words (pair_orbit_sum).  Against the
    reference's loop order (translation, then pair and shift, every term
    located on its own): the same entries, values to rounding, including
    twists (where the Coulomb table loses its |dm| <-> nphi - |dm| symmetry,
    so the image's pair labelling matters), spinful and monolayer sectors."""
    sec = blqh_amd.Sector(*case)
    par = blqh_amd.params(case[1], **kw)
    ids, _ = hostmirror.basis(sec)
    cap = 400 * len(ids) + 64
    rp, col, val = hostmirror.hamil(sec, par, len(ids), cap)
    rp2, col2, val2 = hostmirror.hamil_ql_major(sec, par, len(ids), cap)
    np.testing.assert_array_equal(rp, rp2)
    np.testing.assert_array_equal(col, col2)
    np.testing.assert_allclose(val, val2, rtol=0, atol=1e-13 * np.abs(val2).max())
    assert np.abs(val2.imag).max() > 0 or case[-1] in (0, -1) or case[0] == 1


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


def _members(rng, P, real_frac, multi_frac, max_items):
    nit = rng.integers(1, max_items + 1, size=P)
    item0 = np.concatenate([[0], np.cumsum(nit)])
    nnz = nit * rng.integers(50, 400, size=P)
    npan = np.where(rng.random(P) < multi_frac, rng.integers(2, 5, size=P), 1)
    real = (rng.random(P) < real_frac).astype(np.int32)
    return item0, nnz, npan, real


@pytest.mark.parametrize("P,real_frac,multi_frac,max_items", [
    (1, 0.0, 0.0, 74), (2, 1.0, 1.0, 800), (6, 0.5, 0.0, 30), (144, 0.17, 0.0, 74),
    (64, 0.1, 0.0, 63), (40, 0.5, 0.3, 200), (300, 0.5, 0.05, 9)])
def test_xcd_work_lists(P, real_frac, multi_frac, max_items):
    """The batched K1's per-XCD lists: every item of a kind exactly once,
    ascending within an XCD, several-panel members dealt item i -> XCD i mod 8,
    one-panel members whole when there are many and spread when there are
    few, chunk loads balanced to within the largest chunk."""
    rng = np.random.default_rng(P)
    item0, nnz, npan, real = _members(rng, P, real_frac, multi_frac, max_items)
    items, seg = hostmirror.xcd_lists(item0, nnz, npan, real)
    nitems = int(item0[-1])
    owner = np.repeat(np.arange(P), np.diff(item0))
    assert seg[0] == 0 and seg[26] == 2 * nitems and seg[8] == seg[9] and seg[17] == seg[18] == nitems
    for kind in range(3):
        mine = np.ones(P, bool) if kind == 2 else (real == (1 if kind == 0 else 0))
        got = np.concatenate([items[seg[9 * kind + x]:seg[9 * kind + x + 1]] for x in range(8)])
        np.testing.assert_array_equal(np.sort(got), np.flatnonzero(mine[owner]))
        small = mine & (npan == 1)
        cmax = max(1, -(-int(np.diff(item0)[small].sum()) // 64))
        load = np.zeros(8)
        for x in range(8):
            lst = items[seg[9 * kind + x]:seg[9 * kind + x + 1]]
            assert np.all(np.diff(lst) > 0)
            multi = lst[npan[owner[lst]] > 1]
            assert np.all(multi % 8 == x)
            load[x] = (nnz[owner[lst]] / np.diff(item0)[owner[lst]]).sum()
        if small.any() and not (mine & (npan > 1)).any():
            # greedy largest-first dealing: loads differ by at most one chunk
            wmax = max((nnz[p] / (item0[p + 1] - item0[p])) * min(cmax, item0[p + 1] - item0[p])
                       for p in np.flatnonzero(small))
            assert load.max() - load.min() <= wmax * (1 + 1e-12)
        whole = small & (np.diff(item0) <= cmax)
        for p in np.flatnonzero(whole):
            xs = {x for x in range(8) if np.isin(np.arange(item0[p], item0[p + 1]),
                                                 items[seg[9 * kind + x]:seg[9 * kind + x + 1]]).any()}
            assert len(xs) == 1, (kind, p)
        if small.sum() <= 2 and np.diff(item0)[small].sum() >= 64:
            assert all(seg[9 * kind + x + 1] > seg[9 * kind + x] for x in range(8))


def test_small_modulo_by_fp32_reciprocal():
    """The build's hop-target shift on the device: s mod d from the FP32
    reciprocal of d (the hardware's approximate reciprocal is within 1 ulp;
    checked here for every reciprocal within 2 ulp of 1/d), truncated
    quotient, one correction step either way.  Runs the product's own helper
    (mod_by_rcp in rowterms.hpp, which the device's mod_small calls with the
    hardware reciprocal) through the CPU mirror.  Exact for 0 <= s < 2^16,
    1 <= d <= 64 (the build uses s < 64)."""
    s = np.arange(1 << 16, dtype=np.int64)
    for d in range(1, 65):
        r0 = np.float32(1.0) / np.float32(d)
        up1 = np.nextafter(r0, np.float32(2))
        dn1 = np.nextafter(r0, np.float32(0))
        cands = [r0, up1, np.nextafter(up1, np.float32(2)), dn1, np.nextafter(dn1, np.float32(0))]
        for rc in cands:
            np.testing.assert_array_equal(hostmirror.mod_by_rcp(1 << 16, d, rc), s % d)


def test_mod_by_rcp_fails_with_a_bad_reciprocal():
    """The check above can fail: a reciprocal 1 % off gives wrong remainders."""
    d = 7
    bad = hostmirror.mod_by_rcp(1 << 16, d, np.float32(1.01 / d))
    assert np.any(bad != np.arange(1 << 16) % d)


@pytest.mark.parametrize("case", [
    (2, 8, 8, 4, -1, 4, 4), (2, 8, 8, -1, -1, 3, 2), (2, 10, 10, 5, -1, 1, 7), (2, 9, 6, -1, -1, 2, 1),
    (2, 12, 12, -1, -1, 1, 2), (2, 14, 14, -1, -1, 0, 0), (2, 14, 14, 7, -1, 7, 7), (2, 16, 16, 8, -1, 0, 0),
    (4, 6, 6, 3, 3, 0, 1), (4, 6, 6, -1, -1, 1, 2), (4, 8, 8, 4, 4, 3, 5), (4, 4, 4, 2, 2, 1, 1)],
    ids=lambda c: "_".join(map(str, c)))
def test_hash_order_replay_equals_the_real_map(case):
    """The basis step's representative rule needs the iteration order of the
    reference's std::unordered_map; the library replays the container's node
    list with index arrays (basis.hip hash_iteration_order).  It must equal a
    real map filled the same way, word for word, on every sector the tests
    and the benchmarks use (up to 10.4M words here)."""
    ok, n = hostmirror.hash_order_check(blqh_amd.Sector(*case))
    assert n > 0 and ok, (case, n)


@pytest.mark.parametrize("case", [
    (2, 8, 8, 4, -1, 4, 4), (2, 8, 8, -1, -1, 3, 2), (2, 10, 10, 5, -1, 1, 7), (2, 9, 6, -1, -1, 2, 1),
    (2, 12, 12, -1, -1, 1, 2), (2, 14, 14, -1, -1, 0, 0), (2, 14, 14, 7, -1, 7, 7), (2, 16, 16, 8, -1, 0, 0),
    (4, 6, 6, 3, 3, 0, 1), (4, 6, 6, -1, -1, 1, 2), (4, 8, 8, 4, 4, 3, 5), (4, 4, 4, 2, 2, 1, 1),
    (1, 12, 4, -1, -1, 3, 3), (1, 18, 6, -1, -1, 0, 0), (1, 9, 3, -1, -1, 4, 2), (2, 5, 1, -1, -1, 2, 0)],
    ids=lambda c: "_".join(map(str, c)))
def test_hash_order_levels_equal_the_real_map(case):
    """The device computes the map's iteration order without the list: words
    generated in insertion order from the bit-reversed sector, then list
    positions level by level of the container's rehash schedule (buckets by
    decreasing earliest time, nodes of a bucket by decreasing time).  The
    same steps run sequentially here and must give a real std::unordered_map's
    order, word for word, on spinless, spinful and monolayer sectors."""
    ok, n = hostmirror.hash_order_levels_check(blqh_amd.Sector(*case))
    assert n > 0 and ok, (case, n)


MONO = [(9, 3, 4, 2), (9, 3, 1, -1), (12, 4, 3, 3), (12, 6, 1, 2), (10, 5, 2, 1), (8, 4, 1, 1), (7, 3, 2, 0),
        (18, 6, 3, 3), (11, 4, 5, 0), (14, 4, 0, 1)]


@pytest.mark.parametrize("nphi,nel,J,K", MONO, ids=[f"n{a}_e{b}_J{c}_K{d}" for a, b, c, d in MONO])
def test_mirror_monolayer_matches_oracle(oracle_mod, nphi, nel, J, K):
    """Monolayer mode (nspecies = 1): gcd translations, folded Coulomb table,
    occupied-orbital sign at odd nphi, E_cl background."""
    O = oracle_mod
    gamma = nel / 4.0
    d = O.monolayer(nphi, nel, J, K, gamma=gamma)
    sec = blqh_amd.Sector(1, nphi, nel, -1, -1, J, K)
    ids, per = hostmirror.basis(sec)
    np.testing.assert_array_equal(ids, d["id"])
    np.testing.assert_array_equal(per, d["period"])
    par = blqh_amd.params(nphi, gamma=gamma)
    rp, col, val = hostmirror.hamil(sec, par, len(ids), 2 * len(d["col"]) + 64)
    np.testing.assert_array_equal(rp, d["rowptr"])
    np.testing.assert_array_equal(col, d["col"])
    np.testing.assert_allclose(val, d["val"], rtol=0, atol=1e-13 * np.abs(d["val"]).max())


@pytest.mark.parametrize("m, seed", [(2, 0), (3, 1), (40, 2), (150, 3)])
def test_low_ritz_vectors_match_lapack(m, seed):
    """The batched solver's host Ritz vectors (bisection + pivoted inverse
    iteration) against LAPACK's tridiagonal eigensolver: eigenvectors k = 0..2
    up to sign, and the excited mix = (v1 + v2)/sqrt(2) up to the signs of
    v1 and v2 (unit norm, orthogonal to v0, equal weight on v1 and v2)."""
    from scipy.linalg import eigh_tridiagonal
    rng = np.random.default_rng(seed)
    a = rng.normal(size=m) * 2.0
    b = np.abs(rng.normal(size=m - 1)) + 0.1  # Lanczos off-diagonals are positive
    w, V = eigh_tridiagonal(a, b)
    for k in range(min(3, m)):
        x = hostmirror.tridiag_vector_k(a, b, k)
        assert abs(np.linalg.norm(x) - 1) < 1e-12
        assert abs(abs(x @ V[:, k]) - 1) < 1e-9, (m, k)
    c0 = hostmirror.tridiag_vector_k(a, b, 0)
    xm = hostmirror.excited_mix(a, b, c0)
    assert abs(np.linalg.norm(xm) - 1) < 1e-12
    assert abs(xm @ V[:, 0]) < 1e-9
    if m == 2:
        assert abs(abs(xm @ V[:, 1]) - 1) < 1e-9
    else:
        c = np.abs(V[:, 1:3].T @ xm)
        assert np.allclose(c, 1 / np.sqrt(2), atol=1e-9), c
    assert np.all(hostmirror.excited_mix(a[:1], b[:0], c0[:1]) == 0)
