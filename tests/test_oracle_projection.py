"""The oracle's Hamiltonian pinned by what the mathematics fixes, beyond the
README sector and Hermiticity (oracle/blqh_oracle.cpp follows
BLQH_without_spin/lanczos_hamiltonian.cpp:87-358 and
BLQH_with_spin/lanczos_hamiltonian.cpp:105-908).

1. Bloch projection.  Row i of the reference's (J, K) Hamiltonian sums, over
   the translation images L = T^ql c_i (ql < D_i) and every hop of L, the
   element V * sign * FT[ql][qr] / sqrt(D_l D_r), qr locating the target in
   its orbit (lanczos_hamiltonian.cpp:104-166).  That is
   H[i][j] = <w_j| H_F |w_i> with the truncated Bloch vectors
   w_i = D_i^-1/2 sum_{q < D_i} exp(2 pi i q K / nphi) s(c_i, q) |T^q c_i>
   (s the translation's fermion sign) and H_F the same operator on plain
   Fock words (the reference's K = -1 mode, no translation reduction).  The
   test builds H_F over every word of the electron number, projects it, and
   compares every entry of the (J, K) matrix: complex K, tunnelling, bias,
   Zeeman, spinless and spinful.  A conjugated or transposed phase, a wrong
   1/sqrt(D) or a wrong translation sign fails it.
2. Spin-polarised spinful = spinless.  With every electron spin-up (or every
   one spin-down) the spinful Hamiltonian is the spinless bilayer's on the
   same layer counts: same spectrum.  Pins the same-spin pair blocks.
3. SU(2).  At Delta_Z = 0 and no tunnelling the Coulomb Hamiltonian is spin
   rotation invariant, so every level of the S_z = m + 1 sector is a level
   of the S_z = m sector (m >= 0).  A missing or mis-weighted inter-spin
   pair block breaks it.
4. The printed Lanczos energy of a complex sector: the
   reference prints Re<psi|H psi> with psi = sum_l c_l conj(v_l)
   (lanczos_hamiltonian.cpp:624-632,646-656), i.e. psi = conj(u0), which is
   not E0 when H is complex.  The oracle's converged value equals
   Re(conj(u0)^H H conj(u0)) from LAPACK to 1e-10."""
import numpy as np
import pytest
import scipy.sparse as sp


def _ham_csr(O, sec, par):
    B = O.Basis(sec)
    H = O.Hamiltonian(B, par)
    return B, H, sp.csr_matrix((H.val, H.col, H.rowptr), shape=(H.n, H.n))


PROJ = [
    ("spinless_alldens_complex_tunnel", dict(S=2, nphi=6, nel=6, nel_up=-1, J=1, K=1),
     dict(d=1.0, dsas=0.1, dv=0.05)),
    ("spinless_alldens_K2", dict(S=2, nphi=6, nel=6, nel_up=-1, J=0, K=2), dict(d=0.4, dsas=0.07)),
    ("spinless_split_complex", dict(S=2, nphi=6, nel=6, nel_up=3, J=2, K=1), dict(d=2.0)),
    ("spinless_nphi8_complex_tunnel", dict(S=2, nphi=8, nel=8, nel_up=-1, J=3, K=3),
     dict(d=1.0, dsas=0.05, dv=0.02)),
    ("spinful_fixed_complex", dict(S=4, nphi=4, nel=4, nel_up=2, nel_su=2, J=1, K=1),
     dict(d=1.0, dsas=0.05, dv=0.02, dz=0.01)),
    ("spinful_alldens_complex", dict(S=4, nphi=4, nel=4, nel_up=-1, nel_su=-1, J=0, K=1),
     dict(d=1.5, dsas=0.05, dv=0.02, dz=0.01)),
    ("spinful_alldens_K2", dict(S=4, nphi=4, nel=4, nel_up=-1, nel_su=-1, J=2, K=2),
     dict(d=0.5, dsas=0.1, dz=0.03)),
]


@pytest.mark.parametrize("name,sec,kw", PROJ, ids=[p[0] for p in PROJ])
def test_bloch_projection(oracle_mod, name, sec, kw):
    O = oracle_mod
    nphi = sec["nphi"]
    par = O.params(nphi, **kw)
    sK = O.sector(**sec)
    BK, HK, AK = _ham_csr(O, sK, par)
    sF = O.sector(**dict(sec, J=-1, K=-1))  # every word of the electron (layer, spin) numbers
    BF, HF, AF = _ham_csr(O, sF, par)
    n, nf = len(BK), len(BF)
    rows, cols, vals = [], [], []
    K = sec["K"]
    for i, (c, D) in enumerate(zip(BK.id, BK.period)):
        for q in range(int(D)):
            w, s = (int(c), 1) if q == 0 else O.translate(sK, int(c), q)
            f = BF.find(w)
            assert f >= 0
            rows.append(f)
            cols.append(i)
            vals.append(s * np.exp(2j * np.pi * q * K / nphi) / np.sqrt(float(D)))
    P = sp.csr_matrix((vals, (rows, cols)), shape=(nf, n))
    proj = (P.conj().T @ AF @ P).toarray()
    HKd = AK.toarray()
    scale = np.abs(HKd).max()
    # entries the reference drops (|h| <= 1e-10) are absent from H_K only
    np.testing.assert_allclose(HKd, proj.T, rtol=0, atol=2e-10 + 1e-13 * scale)
    assert np.abs(HKd.imag).max() > 1e-3 or K in (0, nphi // 2)  # the phases are exercised


def _spec(O, sec, par):
    B = O.Basis(sec)
    H = O.Hamiltonian(B, par)
    return np.linalg.eigvalsh(H.dense())


@pytest.mark.parametrize("J,K", [(0, 0), (1, 2), (3, 1)])
def test_spin_polarised_spinful_is_spinless(oracle_mod, J, K):
    O = oracle_mod
    nphi, nel, nu = 6, 6, 3
    par = O.params(nphi, d=1.2)
    ref = _spec(O, O.sector(S=2, nphi=nphi, nel=nel, nel_up=nu, J=J, K=K), par)
    for nsu in (nel, 0):  # all up: species 0 and 2; all down: species 1 and 3
        spin = _spec(O, O.sector(S=4, nphi=nphi, nel=nel, nel_up=nu, nel_su=nsu, J=J, K=K), par)
        assert len(spin) == len(ref)
        np.testing.assert_allclose(spin, ref, rtol=0, atol=1e-11 * np.abs(ref).max())


@pytest.mark.parametrize("J,K", [(0, 0), (1, 1), (2, 3)])
def test_su2_multiplets(oracle_mod, J, K):
    O = oracle_mod
    nphi, nel, nu = 4, 6, 3
    par = O.params(nphi, d=0.8)  # Delta_Z = 0, no tunnelling
    specs = {nsu: _spec(O, O.sector(S=4, nphi=nphi, nel=nel, nel_up=nu, nel_su=nsu, J=J, K=K), par)
             for nsu in (3, 4, 5)}
    scale = np.abs(specs[3]).max()
    for lo, hi in ((3, 4), (4, 5)):
        big, small = specs[lo], specs[hi]
        assert len(small) < len(big)
        miss = np.array([np.abs(big - e).min() for e in small])
        assert miss.max() < 1e-10 * scale, (lo, hi, miss.max())
    # and the containment is not trivial: S_z = 1 has levels (S = 1) the S_z = 2 sector lacks
    extra = np.array([np.abs(specs[5] - e).min() for e in specs[4]])
    assert extra.max() > 1e-6


def test_su2_fails_without_an_interspin_block(oracle_mod):
    """Control for the test above: a Hamiltonian that is not spin rotation
    invariant (here through the Zeeman term, which shifts the S_z sectors
    apart) breaks the containment, so the check has teeth."""
    O = oracle_mod
    par = O.params(4, d=0.8, dz=0.1)
    big = _spec(O, O.sector(S=4, nphi=4, nel=6, nel_up=3, nel_su=3, J=0, K=0), par)
    small = _spec(O, O.sector(S=4, nphi=4, nel=6, nel_up=3, nel_su=4, J=0, K=0), par)
    assert max(np.abs(big - e).min() for e in small) > 1e-3


@pytest.mark.parametrize("J,K", [(1, 1), (3, 2)])
def test_complex_sector_printed_energy(oracle_mod, J, K):
    O = oracle_mod
    nphi = 8
    B = O.Basis(O.sector(S=2, nphi=nphi, nel=8, nel_up=-1, J=J, K=K))
    H = O.Hamiltonian(B, O.params(nphi, d=1.0, dsas=0.05))
    Hd = H.dense()
    w, V = np.linalg.eigh(Hd)
    assert w[1] - w[0] > 1e-6
    u = np.conj(V[:, 0])
    printed = np.real(np.vdot(u, Hd @ u))
    r = H.lanczos(200, seed=2)
    assert r["E_rq"] == pytest.approx(printed, rel=1e-10)
    # for a complex sector the printed value is not the ground-state energy
    assert abs(printed - w[0]) > 1e-6 * abs(w[0])


# ---- the checks above fail on plausible mistakes planted in the oracle ----------
def _projection_error(O, sec, kw):
    nphi = sec["nphi"]
    par = O.params(nphi, **kw)
    sK = O.sector(**sec)
    BK, HK, AK = _ham_csr(O, sK, par)
    BF, HF, AF = _ham_csr(O, O.sector(**dict(sec, J=-1, K=-1)), par)
    rows, cols, vals = [], [], []
    for i, (c, D) in enumerate(zip(BK.id, BK.period)):
        for q in range(int(D)):
            w, s = (int(c), 1) if q == 0 else O.translate(sK, int(c), q)
            rows.append(BF.find(w))
            cols.append(i)
            vals.append(s * np.exp(2j * np.pi * q * sec["K"] / nphi) / np.sqrt(float(D)))
    P = sp.csr_matrix((vals, (rows, cols)), shape=(len(BF), len(BK)))
    return np.abs(AK.toarray() - (P.conj().T @ AF @ P).toarray().T).max()


@pytest.mark.parametrize("plant", [3, 4], ids=["conjugated_phase", "sqrt_Dl_only"])
def test_projection_catches_planted_mistakes(oracle_mod, plant):
    O = oracle_mod
    # a sector with short orbits (period 3 of 6) and complex phases
    sec, kw = dict(S=2, nphi=6, nel=6, nel_up=-1, J=1, K=1), dict(d=1.0, dsas=0.1)
    assert _projection_error(O, sec, kw) < 1e-9
    with O.planted(plant):
        assert _projection_error(O, sec, kw) > 1e-3


@pytest.mark.parametrize("plant", [1, 2], ids=["no_block_01", "wrong_table_23"])
def test_su2_catches_planted_interspin_mistakes(oracle_mod, plant):
    O = oracle_mod
    par = O.params(4, d=0.8)
    sec = dict(S=4, nphi=4, nel=6, nel_up=3, J=0, K=0)
    with O.planted(plant):
        big = _spec(O, O.sector(nel_su=3, **sec), par)
        small = _spec(O, O.sector(nel_su=4, **sec), par)
    assert max(np.abs(big - e).min() for e in small) > 1e-4
