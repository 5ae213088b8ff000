"""Ground-state observables on the GPU against the oracle's restatement of the
reference loops, on the same state (the engine's Lanczos ground state)."""
import numpy as np
import pytest

import blqh_amd
import oracle as O

pytestmark = pytest.mark.gpu

CASES = [
    ("spinless_real", dict(S=2, nphi=6, nel=6, nel_up=-1, nel_su=-1, J=3, K=3), 1.0, 0.05),
    ("spinless_complex", dict(S=2, nphi=6, nel=6, nel_up=-1, nel_su=-1, J=1, K=2), 0.5, 0.3),
    ("spinless_u", dict(S=2, nphi=8, nel=8, nel_up=4, nel_su=-1, J=4, K=4), 3.0, 0.0),
    ("spinless_no_translation", dict(S=2, nphi=6, nel=6, nel_up=-1, nel_su=-1, J=2, K=-1), 1.0, 0.1),
    ("spinful_all", dict(S=4, nphi=4, nel=4, nel_up=-1, nel_su=-1, J=0, K=1), 1.0, 0.05),
    ("spinful_fixed", dict(S=4, nphi=4, nel=4, nel_up=2, nel_su=2, J=1, K=0), 1.0, 0.05),
    ("spinful_polarised", dict(S=4, nphi=4, nel=4, nel_up=-1, nel_su=4, J=2, K=2), 1.0, 0.05),
    ("spinful_zeeman", dict(S=4, nphi=4, nel=4, nel_up=-1, nel_su=-1, J=1, K=3), 1.5, 0.1, 0.03),
]


@pytest.mark.parametrize("name,sec,d,dsas,dz", [c if len(c) == 5 else c + (0.0,) for c in CASES],
                         ids=[c[0] for c in CASES])
def test_observables_match_oracle(name, sec, d, dsas, dz):
    print(name, flush=True)
    B = blqh_amd.Basis(sec["nphi"], sec["nel"], sec["nel_up"], sec["J"], sec["K"], nspecies=sec["S"],
                       nel_su=sec["nel_su"])
    H = blqh_amd.Hamiltonian(B, blqh_amd.params(sec["nphi"], d=d, delta_sas=dsas, delta_z=dz))
    psi = H.lanczos(lam=300, seed=3, check_every=1)["psi"]
    got = blqh_amd.observables(B, psi)
    OB = O.Basis(O.sector(**sec))
    assert np.array_equal(OB.id, B.id)
    ref = O.observables(OB, psi)
    for k in ("pseudospin_sz", "pseudospin_sx", "sz", "upper_sz", "lower_sz", "spinflip"):
        assert got[k] == pytest.approx(ref[k], rel=1e-12, abs=1e-14), k
    np.testing.assert_allclose(got["occupation"], ref["occupation"], rtol=1e-12, atol=1e-14)
