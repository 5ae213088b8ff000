"""Twisted-boundary sweep (the reference's spin-stiffness loop,
BLQH_without_spin/main.cpp:167-181: E0 over theta_y in [0, 2 pi) at fixed
sector and parameters), solved as one batch of Hamiltonians: every point's
ground energy equals dense diagonalisation of the oracle's Hamiltonian."""
import numpy as np
import pytest

import blqh_amd
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("S,args,kw", [
    (2, (8, 8, -1, 0, 0), dict(d=1.0, delta_sas=0.05)),
    (4, (4, 4, -1, 1, 0), dict(d=1.5, delta_sas=0.1, delta_z=0.01)),
], ids=["spinless", "spinful"])
def test_theta_y_sweep_matches_oracle(S, args, kw):
    nphi, nel, nu, J, K = args
    B = blqh_amd.Basis(nphi, nel, nu, J, K, nspecies=S)
    OB = O.Basis(O.sector(S=S, nphi=nphi, nel=nel, nel_up=nu, J=J, K=K))
    thetas = [i / 4.0 * np.pi for i in range(8)]
    hs = [blqh_amd.Hamiltonian(B, blqh_amd.params(nphi, theta_y=t, **kw)) for t in thetas]
    res = blqh_amd.Batch(hs).lanczos(lam=300, seeds=1, check_every=2, want_psi=False)
    m = {"delta_sas": "dsas", "delta_z": "dz"}
    okw = {m.get(k, k): v for k, v in kw.items()}
    e = []
    for t, r in zip(thetas, res):
        w = np.linalg.eigvalsh(O.Hamiltonian(OB, O.params(nphi, theta_y=t, **okw)).dense())
        assert r["converged"] == 1
        assert r["E0"] == pytest.approx(w[0], rel=1e-10), t
        e.append(r["E0"])
    assert max(e) - min(e) > 1e-6  # the twist changes the ground energy
