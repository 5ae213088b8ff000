"""A warm start orthogonal to the ground state: the ground state of the
previous parameter point across a level crossing between symmetry classes of
the sector (spinless bilayer N_phi = 10, all densities, (J, K) = (3, 2),
d = 0.2, Delta_SAS 0.05 -> 0.1: overlap 1e-9).  Lanczos from it converges,
correctly by its own test, to an excited level; start_mix adds the seeded
random unit vector and recovers the ground state."""
import numpy as np
import pytest
import scipy.sparse as sp

import blqh_amd

pytestmark = pytest.mark.gpu


def test_start_mix_recovers_ground_state_across_level_crossing():
    B = blqh_amd.Basis(10, 10, -1, 3, 2)
    H0 = blqh_amd.Hamiltonian(B, blqh_amd.params(10, d=0.2, delta_sas=0.05))
    H1 = blqh_amd.Hamiltonian(B, blqh_amd.params(10, d=0.2, delta_sas=0.1))
    rp, col, val = H1.export()
    w = np.linalg.eigvalsh(sp.csr_matrix((val, col, rp), shape=(H1.n, H1.n)).toarray())
    prev = np.conj(H0.lanczos(lam=300, seed=1)["psi"])  # eigenvector of H0
    trapped = H1.lanczos(lam=300, seed=1, start=prev)
    assert trapped["converged"] == 1 and trapped["E0"] > w[0] + 1e-3  # an excited level
    for mix in (0.1, 0.01):
        r = H1.lanczos(lam=300, seed=1, start=prev, start_mix=mix)
        assert r["converged"] == 1
        assert r["E0"] == pytest.approx(w[0], rel=1e-10), mix
    rb = blqh_amd.Batch([H1]).lanczos(lam=300, seeds=1, starts=[prev], start_mix=0.1, want_psi=False)[0]
    assert rb["E0"] == pytest.approx(w[0], rel=1e-10)
    with pytest.raises(blqh_amd.BlqhError):
        H1.lanczos(lam=300, seed=1, start=prev, start_mix=-1.0)
