"""The oracle's monolayer sectors (oracle/blqh_oracle.cpp, following
Monolayer/basis.cpp:55-289 and Monolayer/lanczos_hamiltonian.cpp:53-176)
pinned by the Monolayer tree's own check data and by the Bloch projection.

1. Ground-state energy per particle over the filling scans of
   Monolayer/check/gs_energy_nel{4,5,6} (35 points, 6 significant digits):
   lx / ly = nel / 4 (Monolayer/main.cpp:32), every J block (the scan builds
   the whole space, Monolayer/main.cpp:172-186), E_gs / nel.
2. The 10 lowest levels of every (J, K) sector at nphi = 18, nel = 6,
   lx / ly = 1 (Monolayer/check/ek_nphi_18_nel_6, 108 rows, rows j-major,
   first column |K| = sqrt((j % 6 - 3)^2 + (k % 6 - 3)^2) sqrt(2 pi / 18)).
   These pin the K machinery (gcd translations, representatives, periods,
   phases), including the states of short orbits that a K sector keeps even
   when the orbit carries no momentum K (the tree keeps every orbit).
3. Bloch projection: entry (i, j) of the (J, K) matrix (row i = the source
   state, as the tree fills it) equals <w_j| H_F |w_i>, H_F the K < 0 matrix over every word of the electron number and
   w_i = D_i^-1/2 sum_{q < D_i} exp(2 pi i q K / C) s(c_i, q) |T^q c_i>
   (T a rotation by nphi / C orbitals, C = gcd(nel, nphi)).
4. Fermion signs: the K < 0 matrix equals the second-quantised operator
   sum_{n<m} sum_t 2 V(|n - mt|, |t|) c+_nt c+_mt c_m c_n + E_cl nel built
   with Jordan-Wigner signs on every word, at odd nphi, where a hop with one
   wrapped electron distinguishes the occupied-orbital count from the
   bilayer trees' empty-orbital count.  (The check data cannot: at odd nphi
   that difference maps J blocks onto each other and the scans take the
   minimum over J.)  V comes from its definition, whose values the check
   data pin.
5. Planted mistakes: the empty-orbital sign fails 4, a conjugated phase
   fails 3."""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

CHECK = os.path.join(os.path.dirname(__file__), "golden", "monolayer")


def _csr(d):
    n = len(d["id"])
    return sp.csr_matrix((d["val"], d["col"], d["rowptr"]), shape=(n, n))


def _lowest(A, k=1):
    if A.shape[0] <= 600:
        return np.linalg.eigvalsh(A.toarray())[:k]
    return np.sort(sla.eigsh(A, k=k, which="SA", tol=1e-12)[0])


def _gs_per_particle(O, nphi, nel):
    return min(_lowest(_csr(O.monolayer(nphi, nel, J, -1, gamma=nel / 4.0)))[0] for J in range(nphi)) / nel


@pytest.mark.parametrize("nel", [4, 5, 6])
def test_ground_energy_scan_matches_check_data(oracle_mod, nel):
    data = np.loadtxt(os.path.join(CHECK, f"gs_energy_nel{nel}"))
    for nu, e in data:
        nphi = int(round(nel / nu))
        assert float(f"{nu:.6g}") == pytest.approx(nel / nphi, abs=1e-6)
        got = _gs_per_particle(oracle_mod, nphi, nel)
        assert float(f"{got:.6g}") == e, (nel, nphi, got, e)


def _ek_row(O, r, nphi=18, nel=6):
    j, k = divmod(r, 6)
    A = _csr(O.monolayer(nphi, nel, j, k, gamma=1.0)).toarray()
    assert np.abs(A - A.conj().T).max() < 1e-13
    return j, k, np.linalg.eigvalsh(A)[:10]


def test_k_resolved_spectrum_matches_check_data(oracle_mod):
    data = np.loadtxt(os.path.join(CHECK, "ek_nphi_18_nel_6"))
    assert data.shape == (108, 11)
    worst = 0.0
    for r in range(108):
        j, k, w = _ek_row(oracle_mod, r)
        kk = np.hypot(j % 6 - 3, k % 6 - 3) * np.sqrt(2 * np.pi / 18)
        assert kk == pytest.approx(data[r, 0], abs=1e-5)
        worst = max(worst, np.abs(w - data[r, 1:]).max())
    # 6 significant digits of numbers of magnitude 2.2-2.5: half a unit in the last place
    assert worst <= 5e-6 + 1e-9, worst


def _translate(c, r, nphi, nel):
    """Left rotation by r orbitals, sign (-1)^(w (nel - w)), w = wrapping electrons."""
    m = (1 << nphi) - 1
    w = bin(c >> (nphi - r)).count("1") if r else 0
    return ((c << r) | (c >> (nphi - r))) & m if r else c, (-1) ** (w * (nel - w))


PROJ = [(9, 3, 1, 1), (9, 3, 4, 2), (12, 4, 2, 1), (12, 4, 3, 3), (10, 4, 1, 1), (8, 4, 1, 1), (12, 6, 3, 5),
        (10, 5, 2, 1)]


def _projection_error(O, nphi, nel, J, K):
    C = int(np.gcd(nphi, nel))
    q = nphi // C
    dk = O.monolayer(nphi, nel, J, K)
    dF = O.monolayer(nphi, nel, -1, -1)
    row = {int(w): i for i, w in enumerate(dF["id"])}
    rows, cols, vals = [], [], []
    for i, (c, D) in enumerate(zip(dk["id"], dk["period"])):
        for t in range(int(D)):
            w, s = _translate(int(c), t * q, nphi, nel)
            rows.append(row[w])
            cols.append(i)
            vals.append(s * np.exp(2j * np.pi * t * K / C) / np.sqrt(float(D)))
    P = sp.csr_matrix((vals, (rows, cols)), shape=(len(dF["id"]), len(dk["id"])))
    proj = (P.conj().T @ _csr(dF) @ P).toarray().T
    HK = _csr(dk).toarray()
    # entries the tree drops (|h| <= 1e-10) are absent from H_K only
    err = np.maximum(np.abs(proj - HK) - 2e-10, 0).max() / np.abs(HK).max()
    return err, np.abs(HK.imag).max()


@pytest.mark.parametrize("nphi,nel,J,K", PROJ, ids=[f"n{a}_e{b}_J{c}_K{d}" for a, b, c, d in PROJ])
def test_bloch_projection(oracle_mod, nphi, nel, J, K):
    err, imag = _projection_error(oracle_mod, nphi, nel, J, K)
    assert err < 1e-13
    C = np.gcd(nphi, nel)
    assert imag > 1e-3 or (2 * K) % C == 0  # the phases are exercised


def _jw(c, k, create):
    """c+_k (create) or c_k on word c with the Jordan-Wigner sign of the
    occupied orbitals below k; None when the result vanishes."""
    if ((c >> k) & 1) == create:
        return None, 0
    return c ^ (1 << k), (-1) ** bin(c & ((1 << k) - 1)).count("1")


def _second_quantised(nphi, nel, lx, ly, ecl):
    def V(kx, ky):
        q = np.sqrt(kx * kx / (lx * lx) + ky * ky / (ly * ly)) * 2 * np.pi
        return 2 * np.pi / (q + 1e-30) * np.exp(-q * q / 2)
    Vt = np.zeros((nphi, nphi))
    for s_ in range(nphi):
        for ky in range(nphi):
            Vt[s_, ky] = sum(V(kx, ky) * np.cos(2 * np.pi * s_ * kx / nphi) / (2 * lx * ly)
                             for kx in range(-(nphi // 2), nphi // 2 + 1) if (kx, ky) != (0, 0))
    words = [w for w in range(1 << nphi) if bin(w).count("1") == nel]
    idx = {w: i for i, w in enumerate(words)}
    H = np.zeros((len(words), len(words)))
    for i, c in enumerate(words):
        H[i, i] += ecl * nel
        for n in range(nphi - 1):
            for m in range(n + 1, nphi):
                for t in range(-(nphi // 2), nphi // 2 + 1):
                    nt, mt = (n + t) % nphi, (m - t) % nphi
                    x, sg = c, 1
                    for k, cr in ((n, 0), (m, 0), (mt, 1), (nt, 1)):  # c+_nt c+_mt c_m c_n
                        if x is None:
                            break
                        x, s1 = _jw(x, k, cr)
                        sg *= s1
                    if x is not None:
                        H[i, idx[x]] += 2 * Vt[abs(n - mt), abs(t)] * sg
    return H


@pytest.mark.parametrize("nphi,nel", [(7, 3), (9, 4), (8, 3)])
def test_fermion_signs_are_jordan_wigner(oracle_mod, nphi, nel):
    d = oracle_mod.monolayer(nphi, nel, -1, -1)
    Hq = _second_quantised(nphi, nel, d["lx"], d["ly"], d["ecl"])
    assert [int(w) for w in d["id"]] == [w for w in range(1 << nphi) if bin(w).count("1") == nel]
    Ho = _csr(d).toarray()
    assert np.abs(Ho.imag).max() == 0
    np.testing.assert_allclose(Ho.real, Hq, rtol=0, atol=2e-10 + 1e-13 * np.abs(Hq).max())


def test_planted_mistakes_fail_the_checks(oracle_mod):
    O = oracle_mod
    d = O.monolayer(7, 3, -1, -1)
    Hq = _second_quantised(7, 3, d["lx"], d["ly"], d["ecl"])
    with O.planted(5):
        assert np.abs(_csr(O.monolayer(7, 3, -1, -1)).toarray().real - Hq).max() > 1e-3
    with O.planted(6):
        # a conjugated phase maps K to -K, whose levels are the same: only the
        # entry-by-entry projection sees it
        assert _projection_error(O, 9, 3, 4, 2)[0] > 1e-3
