"""Pass 3 on its own (eigenstates_reconstruction): the ground state and the
first excited state rebuilt in one replay, first_excited_state_energy, and
pass 3 resumed from a Lanczos record read back from its text file."""
import numpy as np
import pytest

import blqh_amd
import oracle as O

pytestmark = pytest.mark.gpu


def _dense(H):
    import scipy.sparse as sp
    rp, col, val = H.export()
    return sp.csr_matrix((val, col, rp), shape=(H.n, H.n)).toarray()


def _ham(sec, d, dsas, dz=0.0):
    B = blqh_amd.Basis(sec["nphi"], sec["nel"], sec["nel_up"], sec["J"], sec["K"], nspecies=sec["S"],
                       nel_su=sec["nel_su"])
    return B, blqh_amd.Hamiltonian(B, blqh_amd.params(sec["nphi"], d=d, delta_sas=dsas, delta_z=dz))


README = dict(S=2, nphi=8, nel=8, nel_up=4, nel_su=-1, J=4, K=4)
SPINFUL = dict(S=4, nphi=4, nel=4, nel_up=-1, nel_su=-1, J=1, K=3)
COMPLEX = dict(S=2, nphi=8, nel=8, nel_up=-1, nel_su=-1, J=0, K=1)
SPINFUL_REAL = dict(S=4, nphi=4, nel=4, nel_up=-1, nel_su=-1, J=1, K=0)


@pytest.mark.parametrize("sec,d,dsas", [(README, 3.0, 0.0), (SPINFUL, 1.5, 0.1), (COMPLEX, 1.0, 0.05)],
                         ids=["readme", "spinful", "complex"])
def test_replayed_ground_state_equals_lanczos_ground_state(sec, d, dsas):
    """Skipping pass 3 and replaying it with the tridiagonal's eigenvector 0
    gives the state blqh_lanczos reconstructs itself, and the replay
    regenerates alpha and beta bit for bit."""
    print(sec, flush=True)
    B, H = _ham(sec, d, dsas)
    full = H.lanczos(lam=200, seed=2, tol=1e-10, store_basis=0)
    r = H.lanczos(lam=200, seed=2, tol=1e-10, want_psi=False, skip_states=True)
    assert np.isnan(r["E_rq"]) and r["passes"] == 1
    np.testing.assert_array_equal(r["alpha"], full["alpha"])
    v = blqh_amd.tridiag_vectors(r["alpha"], r["beta"], 0, 1)
    psi, erq, dev = H.reconstruct(v, r["alpha"], r["beta"], seed=2)
    assert dev == 0.0
    assert abs(np.vdot(psi[0], full["psi"])) == pytest.approx(1.0, abs=1e-12)
    assert erq[0] == pytest.approx(full["E_rq"], rel=1e-12)


# Real sectors (K = 0 or nphi/2), where psi = sum_l c_l conj(v_l) is an
# eigenvector of H itself; lambda inside the window where the first excited
# level has converged and no partly converged ghost of E0 lies more than 1e-6
# above it (without reorthogonalisation a ghost of E0 appears at lambda 34-36
# and about 60 on the README sector, at about 76 on the spinful one)
@pytest.mark.parametrize("sec,d,dsas,lam", [(README, 3.0, 0.0, 42), (SPINFUL_REAL, 1.5, 0.1, 64)],
                         ids=["readme", "spinful"])
def test_first_excited_state(sec, d, dsas, lam):
    """psir_1 with the gap rule: orthogonal to psir_0, a converged eigenpair of
    H, and its energy is the dense spectrum's next distinct level; with the
    reference's own selection (Krylov eigenvector 1) E1_rq is the Rayleigh
    quotient of that Ritz vector."""
    print(sec, flush=True)
    B, H = _ham(sec, d, dsas)
    Hd = _dense(H)
    w = np.linalg.eigvalsh(Hd)
    lev = blqh_amd.spectrum_distinct(w, kmax=2, gap=1e-6)
    r = H.eigenstates(lam=lam, seed=1, tol=0.0, e1_gap=1e-6)
    p0, p1 = r["psi"], r["psi1"]
    assert abs(np.vdot(p0, p1)) < 1e-8
    assert r["E_rq"] == pytest.approx(w[0], rel=1e-10)
    assert r["E1_rq"] == pytest.approx(lev[1], rel=1e-8)
    assert np.linalg.norm(Hd @ p1 - r["E1_rq"] * p1) < 1e-6
    # reference selection: eigenvector 1 of the tridiagonal, whatever it is
    r2 = H.eigenstates(lam=40, seed=1, tol=0.0)
    assert r2["e1_index"] == 1
    p1 = r2["psi1"]
    assert r2["E1_rq"] == pytest.approx(np.vdot(p1, Hd @ p1).real, rel=1e-12)


def test_pass3_resumed_from_record(tmp_path):
    """save_to_file after passes 1-2, read_from_file, then pass 3 from the
    record's seed, alpha, beta and psi_0 (8 significant digits): the state is
    the one of the uninterrupted run."""
    B, H = _ham(README, 3.0, 0.0)
    full = H.lanczos(lam=30, seed=5, tol=0.0, store_basis=0)
    eig, first, vec0 = blqh_amd.tridiag_spectrum(full["alpha"], full["beta"])
    p = tmp_path / "lanczos.rec"
    blqh_amd.record_save(p, H.n, 5, eig, full["beta"], full["alpha"], vec0, first)
    rec = blqh_amd.record_load(p)
    assert rec["nhilbert"] == H.n and rec["lambda"] == 30 and rec["seed"] == 5
    psi, erq, dev = H.reconstruct(rec["psi_0"], rec["overlap"], rec["norm"], seed=rec["seed"])
    assert 0 < dev < 1e-7
    assert abs(np.vdot(psi[0], full["psi"])) == pytest.approx(1.0, abs=1e-12)
    assert erq[0] == pytest.approx(full["E_rq"], rel=1e-12)
    # a record of another run (other seed) does not replay
    _, _, dev_bad = H.reconstruct(rec["psi_0"], rec["overlap"], rec["norm"], seed=6)
    assert dev_bad > 1e-3


def test_reconstruct_many_states_and_start_vector():
    """Several coefficient vectors in one replay equal one replay each, also
    from a caller's start vector."""
    B, H = _ham(COMPLEX, 1.0, 0.05)
    start = np.random.default_rng(0).normal(size=(H.n, 2)) @ np.array([1, 1j])
    r = H.lanczos(lam=50, tol=0.0, start=start, want_psi=False, skip_states=True)
    V = blqh_amd.tridiag_vectors(r["alpha"], r["beta"], 0, 5)
    psi, erq, dev = H.reconstruct(V, r["alpha"], r["beta"], start=start)
    assert dev == 0.0
    for s in range(5):
        p1, e1, _ = H.reconstruct(V[s], start=start)
        assert abs(np.vdot(p1[0], psi[s])) == pytest.approx(1.0, abs=1e-13)
        assert e1[0] == pytest.approx(erq[s], rel=1e-13)
    # Ritz values: the Rayleigh quotients of the states are bounded by the spectrum
    w = np.linalg.eigvalsh(_dense(H))
    assert np.all(erq >= w[0] - 1e-9) and np.all(erq <= w[-1] + 1e-9)


def test_reconstruct_rejects_bad_arguments():
    B, H = _ham(README, 3.0, 0.0)
    with pytest.raises(ValueError):  # 17 states: refused by the binding before the call
        H.reconstruct(np.zeros((17, 4)) + 1.0)
    with pytest.raises(blqh_amd.BlqhError):
        H.reconstruct(np.zeros((1, 4)))  # zero state
    with pytest.raises(blqh_amd.BlqhError):
        H.reconstruct([[1.0, np.nan]])


def test_excited_state_observables_of_oracle_ground_state():
    """Observables of psir_0 from the replay equal the oracle's on the same state."""
    B, H = _ham(SPINFUL, 1.5, 0.1)
    r = H.eigenstates(lam=58, seed=1, tol=0.0, e1_gap=1e-6)
    OB = O.Basis(O.sector(**SPINFUL))
    for psi in (r["psi"], r["psi1"]):
        got, ref = blqh_amd.observables(B, psi), O.observables(OB, psi)
        for k in ("pseudospin_sx", "spinflip", "sz"):
            assert got[k] == pytest.approx(ref[k], rel=1e-12, abs=1e-14), k
