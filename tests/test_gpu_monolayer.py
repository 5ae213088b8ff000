"""Monolayer mode (nspecies = 1) of the engine on the GPU, against the
oracle's monolayer sectors (Monolayer/basis.cpp:55-289,
Monolayer/lanczos_hamiltonian.cpp:53-176) and the Monolayer tree's check
data (tests/golden/monolayer: Monolayer/check/*).

1. Basis ids and periods bit-exact, CSR rows and columns exact, values to
   1e-13, for real and complex (J, K) sectors, odd and even nphi, and K < 0.
2. The filling scans: for each of the 35 (nel, nphi) points every J block is
   solved in one batch; min E0 / nel equals the check value at 6
   significant digits (lx / ly = nel / 4).
3. The K-resolved levels at nphi = 18, nel = 6, lx / ly = 1: the ground
   level of all 108 (J, K) sectors (one batch) within half a unit of the 6th
   digit; the single solver and the dense diagonalisation agree on the
   check file's |K| = 0 row (E0 = -2.48561).
4. Arguments a monolayer does not take are refused."""
import os

import numpy as np
import pytest

import blqh_amd

pytestmark = pytest.mark.gpu

CHECK = os.path.join(os.path.dirname(__file__), "golden", "monolayer")

SECTORS = [(9, 3, 4, 2), (9, 4, 1, -1), (12, 4, 3, 3), (12, 6, 1, 2), (10, 5, 2, 1), (18, 6, 3, 3),
           (11, 4, 5, 0), (15, 5, 0, 4)]


@pytest.mark.parametrize("nphi,nel,J,K", SECTORS, ids=[f"n{a}_e{b}_J{c}_K{d}" for a, b, c, d in SECTORS])
def test_monolayer_sector_parity(oracle_mod, nphi, nel, J, K):
    print(f"monolayer nphi={nphi} nel={nel} J={J} K={K}", flush=True)
    gamma = nel / 4.0
    d = oracle_mod.monolayer(nphi, nel, J, K, gamma=gamma)
    B = blqh_amd.Basis(nphi, nel, -1, J, K, nspecies=1)
    np.testing.assert_array_equal(B.id, d["id"])
    np.testing.assert_array_equal(B.period, d["period"])
    H = blqh_amd.Hamiltonian(B, blqh_amd.params(nphi, gamma=gamma))
    rp, col, val = H.export()
    np.testing.assert_array_equal(rp, d["rowptr"])
    np.testing.assert_array_equal(col.astype(np.int64), d["col"])
    np.testing.assert_allclose(val, d["val"], rtol=0, atol=1e-13 * np.abs(d["val"]).max())
    w = np.linalg.eigvalsh(_dense(d))
    r = H.lanczos(lam=200, seed=3, tol=1e-12)
    assert r["converged"] == 1
    assert r["E0"] == pytest.approx(w[0], rel=1e-10)


def _dense(d):
    n = len(d["id"])
    A = np.zeros((n, n), np.complex128)
    for i in range(n):
        A[i, d["col"][d["rowptr"][i]:d["rowptr"][i + 1]]] = d["val"][d["rowptr"][i]:d["rowptr"][i + 1]]
    return A


@pytest.mark.parametrize("nel", [4, 5, 6])
def test_filling_scan_matches_check_data(nel):
    data = np.loadtxt(os.path.join(CHECK, f"gs_energy_nel{nel}"))
    for nu, e in data:
        nphi = int(round(nel / nu))
        print(f"scan nel={nel} nphi={nphi}", flush=True)
        par = blqh_amd.params(nphi, gamma=nel / 4.0)
        hams = [blqh_amd.Hamiltonian(blqh_amd.Basis(nphi, nel, -1, J, -1, nspecies=1), par) for J in range(nphi)]
        bt = blqh_amd.Batch(hams)
        res = bt.lanczos(lam=300, seeds=[7 + J for J in range(nphi)], tol=1e-12, check_every=2, want_psi=False)
        bt.close()
        assert all(x["converged"] == 1 for x in res)
        got = min(x["E0"] for x in res) / nel
        assert float(f"{got:.6g}") == e, (nel, nphi, got, e)


def test_k_resolved_ground_levels_match_check_data():
    data = np.loadtxt(os.path.join(CHECK, "ek_nphi_18_nel_6"))
    par = blqh_amd.params(18, gamma=1.0)
    hams = []
    for r in range(108):
        j, k = divmod(r, 6)
        hams.append(blqh_amd.Hamiltonian(blqh_amd.Basis(18, 6, -1, j, k, nspecies=1), par))
    print("built 108 sectors", flush=True)
    bt = blqh_amd.Batch(hams)
    res = bt.lanczos(lam=200, seeds=list(range(108)), tol=1e-12, check_every=2, want_psi=False)
    bt.close()
    e0 = np.array([x["E0"] for x in res])
    assert all(x["converged"] == 1 for x in res)
    assert np.abs(e0 - data[:, 1]).max() <= 5e-6 + 1e-9
    # the |K| = 0 row (j = 3, k = 3): single solver and dense diagonalisation
    H = hams[21]
    assert data[21, 0] == 0.0
    s = H.lanczos(lam=200, seed=1, tol=1e-12)
    e_dense, _, _ = H.dense_ground_state()
    assert s["E0"] == pytest.approx(e_dense, rel=1e-10)
    assert float(f"{e_dense:.6g}") == data[21, 1] == -2.48561


def test_monolayer_refuses_bilayer_arguments():
    with pytest.raises(blqh_amd.BlqhError):
        blqh_amd.Basis(12, 4, 2, 0, -1, nspecies=1)  # layer occupation
    with pytest.raises(blqh_amd.BlqhError):
        blqh_amd.Basis(12, 4, -1, 0, 4, nspecies=1)  # K >= gcd(4, 12)
    B = blqh_amd.Basis(12, 4, -1, 0, 1, nspecies=1)
    with pytest.raises(blqh_amd.BlqhError):
        blqh_amd.Hamiltonian(B, blqh_amd.params(12, d=1.0))
    with pytest.raises(blqh_amd.BlqhError):
        blqh_amd.Hamiltonian(B, blqh_amd.params(12, theta_y=0.3))
    H = blqh_amd.Hamiltonian(B, blqh_amd.params(12))
    r = H.lanczos(lam=100, seed=1, tol=1e-12)
    with pytest.raises(blqh_amd.BlqhError):
        blqh_amd.observables(B, r["psi"])
