"""The oracle (CPU restatement of the reference path) pinned against the
reference's own known answers and against checks the mathematics fixes.

Golden values: README.md:58-75 and README.md:85-102 of qiongz/BLQH (dimension,
ED and Lanczos energies per electron, ground-state amplitudes).
"""
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "readme_runs.json")


def golden():
    with open(GOLDEN) as f:
        return json.load(f)


@pytest.mark.parametrize("run", ["d0", "d3"])
def test_readme_runs(oracle_mod, run):
    O = oracle_mod
    g = golden()[run]
    sec = O.sector(S=2, nphi=8, nel=8, nel_up=4, J=4, K=4)
    B = O.Basis(sec)
    assert len(B) == g["nHilbert"]
    H = O.Hamiltonian(B, O.params(8, d=g["d"]))
    w = np.linalg.eigvalsh(H.dense())
    # ED line printed with 6 significant digits
    assert float("%.6g" % (w[0] / 8)) == pytest.approx(g["ed_e_per_el"], abs=0)
    r = H.lanczos(200, seed=1)
    # Lanczos line printed with 10 significant digits
    assert float("%.10g" % (r["E_rq"] / 8)) == g["lanczos_e_per_el"]
    amps = np.sort(np.abs(r["psir"]))[::-1]
    np.testing.assert_allclose(amps[:len(g["amplitudes"])], g["amplitudes"], rtol=0, atol=6e-11)


def test_readme_d0_closed_form(oracle_mod):
    """d=0: pseudospin SU(2) makes the Sz=0 ground state degenerate with the
    filled-upper-layer determinant (a 1x1 sector), and the amplitudes are
    sqrt(k/35)."""
    O = oracle_mod
    par = O.params(8, d=0.0)
    one = O.Hamiltonian(O.Basis(O.sector(S=2, nphi=8, nel=8, nel_up=8, J=4, K=4)), par)
    assert one.n == 1
    H = O.Hamiltonian(O.Basis(O.sector(S=2, nphi=8, nel=8, nel_up=4, J=4, K=4)), par)
    w = np.linalg.eigvalsh(H.dense())
    assert w[0] == pytest.approx(one.val[0].real, rel=1e-12)
    r = H.lanczos(200, seed=3)
    amps = np.sort(np.abs(r["psir"]))[::-1][:10]
    np.testing.assert_allclose(amps, np.sqrt(np.array([4] * 8 + [2, 1]) / 35.0), atol=1e-10)


def test_interlayer_table_equals_intralayer_at_d0(oracle_mod):
    O = oracle_mod
    sec = O.sector(S=2, nphi=8, nel=8, nel_up=4, J=4, K=4)
    C, ec, ft = O.tables(sec, O.params(8, d=0.0))
    np.testing.assert_allclose(C[1], C[0], rtol=0, atol=1e-12)
    # K=4 of nphi=8: FT = exp(i pi (a-b)) = +-1
    np.testing.assert_allclose(ft.imag, 0, atol=1e-14)


@pytest.mark.parametrize("case", [
    (2, 8, 8, 4, -1, 4, 4), (2, 8, 8, -1, -1, 3, 2), (2, 10, 10, 5, -1, 1, 7), (2, 12, 12, 6, -1, 0, 0),
    (2, 9, 6, -1, -1, 2, 1), (2, 8, 8, -1, -1, -1, -1), (4, 6, 6, 3, 3, 0, 1), (4, 6, 6, -1, -1, 1, 2),
    (4, 8, 8, 4, 4, 0, 0)])
def test_basis_matches_reference_build(oracle_mod, case):
    """oracle/_ref holds the reference's own basis.cpp compiled from its sources."""
    O = oracle_mod
    sec = O.sector(*case)
    if not O.ref_available(sec.S == 4):
        pytest.skip("reference basis build not available (needs the reference tree at build time)")
    B = O.Basis(sec)
    rid, rper = O.ref_basis(sec)
    np.testing.assert_array_equal(B.id, rid)
    np.testing.assert_array_equal(B.period, rper)


@pytest.mark.parametrize("spinful", [False, True])
def test_bit_primitives_match_reference(oracle_mod, spinful):
    O = oracle_mod
    if not O.ref_available(spinful):
        pytest.skip("reference basis build not available")
    L = O.ref_lib(spinful)
    S, nphi, nel = (4, 6, 8) if spinful else (2, 10, 9)
    L.ref_prepare(nphi, nel, 4 if spinful else -1, 4 if spinful else -1, 0, 0)
    sec = O.sector(S=S, nphi=nphi, nel=nel)
    rng = np.random.default_rng(7)
    import ctypes as C
    for _ in range(400):
        pos = rng.choice(S * nphi, nel, replace=False)
        c = int(sum(1 << int(p) for p in pos))
        for k in range(1, nphi):
            s1 = C.c_int()
            r1 = L.ref_translate(c, k, C.byref(s1))
            assert O.translate(sec, c, k) == (r1, s1.value)
            r2 = L.ref_inv_translate(c, k, C.byref(s1))
            assert O.inv_translate(sec, c, k) == (r2, s1.value)
        occ = [int(p) for p in sorted(pos)]
        n, m = occ[0], occ[-1]
        for nt in range(S * nphi):
            for mt in range(0, S * nphi, 3):
                assert O.sign2(c, n, m, nt, mt) == L.ref_sign2(c, n, m, nt, mt)
            assert O.sign1(c, n, nt) == L.ref_sign1(c, n, nt)


def test_sector_dimensions_burnside(oracle_mod):
    """Orbit counts agree with Burnside's lemma over the translation group
    restricted to the J sector (a closed form independent of the hash order)."""
    O = oracle_mod
    from itertools import combinations
    nphi, nel = 6, 6
    for J in range(nphi):
        sec = O.sector(S=2, nphi=nphi, nel=nel, nel_up=3, J=J, K=0)
        words = O.sector_configs(sec)
        wset = set(int(w) for w in words)
        # count orbits directly: canonical = min over rotations present in the sector
        canon = set()
        for w in wset:
            rots = [O.inv_translate(sec, w, k)[0] for k in range(nphi)]
            canon.add(min(r for r in rots if r in wset))
        assert len(O.Basis(sec)) == len(canon)
        assert len(words) == len(wset)
    # config sizes quoted for the workloads (all-density and -u) at N_phi=8
    assert len(O.Basis(O.sector(S=2, nphi=8, nel=8, nel_up=4, J=4, K=0))) == 79


def test_start_vector_is_textbook_mt19937(oracle_mod):
    """std::mt19937 is standard-mandated; numpy's legacy RandomState(seed)
    seeds MT19937 the same way (init_genrand).  Entry i = (draw 2i+1, draw 2i)."""
    O = oracle_mod
    n, seed = 64, 12345
    rs = np.random.RandomState(seed)
    raw = rs.randint(0, 2 ** 32, size=2 * n, dtype=np.uint64).astype(np.float64)
    z = (raw[1::2] / 4294967295.0 - 0.5) + 1j * (raw[0::2] / 4294967295.0 - 0.5)
    z /= np.sqrt(np.sum(np.abs(z) ** 2))
    np.testing.assert_allclose(O.start_vector(n, seed), z, rtol=0, atol=1e-15)


def test_mt19937_10000th_draw():
    rs = np.random.RandomState(5489)
    assert int(rs.randint(0, 2 ** 32, size=10000, dtype=np.uint64)[-1]) == 4123659995


@pytest.mark.parametrize("case,kw", [
    ((2, 8, 8, 4, -1, 4, 1), dict(d=1.0)),            # complex sector
    ((2, 8, 8, -1, -1, 2, 3), dict(d=0.5, dsas=0.1)),  # complex, tunnelling
    ((2, 10, 10, 5, -1, 0, 0), dict(d=2.0)),
    ((4, 6, 6, 3, 3, 0, 1), dict(d=1.0, dsas=0.05, dv=0.02, dz=0.01)),
])
def test_oracle_lanczos_vs_lapack(oracle_mod, case, kw):
    """Reference three-pass Lanczos (fixed 200 steps) against LAPACK dense
    diagonalisation of the same matrix; and the reference's printed quantity
    Re<psi|H psi> with psi = sum c_l conj(v_l) equals conj(u0)^T H conj(u0)."""
    O = oracle_mod
    B = O.Basis(O.sector(*case))
    H = O.Hamiltonian(B, O.params(case[1], **kw))
    Hd = H.dense()
    assert np.abs(Hd - Hd.conj().T).max() < 1e-14
    w, V = np.linalg.eigh(Hd)
    r = H.lanczos(200, seed=11)
    assert r["eig"][0] == pytest.approx(w[0], rel=1e-10)
    u0 = V[:, 0]
    printed = np.real(np.conj(u0) @ Hd.T @ u0) if (w[1] - w[0]) > 1e-8 else None
    if printed is not None:
        assert r["E_rq"] == pytest.approx(printed, rel=1e-9)
        ov = abs(np.vdot(np.conj(u0), r["psir"]))
        assert ov > 1 - 1e-9


def test_oracle_tridiag_eigh_vs_numpy(oracle_mod):
    O = oracle_mod
    rng = np.random.default_rng(3)
    for m in (1, 2, 7, 60):
        a = rng.normal(size=m)
        b = rng.normal(size=max(m - 1, 0))
        w, V = O.tridiag_eigh(a, b)
        T = np.diag(a) + np.diag(b, 1) + np.diag(b, -1)
        np.testing.assert_allclose(w, np.linalg.eigvalsh(T), atol=1e-12)
        np.testing.assert_allclose(T @ V, V * w, atol=1e-10)
