"""Learned warm start for the spinful bilayer (32 occupation bits): the
network trained on the GPU on ground states the engine computed for real
spinful sectors (N_phi = 8, N_e = 8, 4 in the upper layer, 4 spin-up: the
reference's `-U 4 -u 4` sectors, 16,024-16,046 states) at d in
{0.5, 1, 1.5, 2, 2.5} supplies start vectors at d between them.  The warm
solves reach the plain solves' energy and need fewer Lanczos steps.  Complex
sectors (K other than 0 and N_phi/2) are not warm-started: the network
predicts a sign, not a phase."""
import numpy as np
import pytest

import blqh_amd

pytestmark = pytest.mark.gpu

NPHI, NEL, NU, NSU = 8, 8, 4, 4
PAR = dict(delta_sas=0.05, delta_v=0.02, delta_z=0.01)
SECTORS = [(0, 0), (3, 4), (5, 0)]  # real: K = 0 or nphi / 2


@pytest.fixture(scope="module")
def trained():
    bases = {s: blqh_amd.Basis(NPHI, NEL, NU, *s, nspecies=4, nel_su=NSU) for s in SECTORS}
    sets = []
    for s in SECTORS:
        for d in (0.5, 1.0, 1.5, 2.0, 2.5):
            H = blqh_amd.Hamiltonian(bases[s], blqh_amd.params(NPHI, d=d, **PAR))
            r = H.lanczos(lam=300, seed=1, check_every=5)
            assert r["converged"] == 1
            sets.append((bases[s], d, PAR["delta_sas"], r["psi"]))
    net = blqh_amd.Mlp(4 * NPHI, seed=7)
    info = net.train(sets, steps=4000, batch=4096, lr=3e-3, lr_final=1e-4, seed=11)
    print("training", info, flush=True)
    return bases, net, info


def test_spinful_training_lowers_the_loss(trained):
    bases, net, info = trained
    assert info["loss_last"] < 0.3 * info["loss_first"]


def test_spinful_warm_start_saves_iterations_same_energy(trained):
    bases, net, info = trained
    plain_it, warm_it = [], []
    for s, B in bases.items():
        for d in (0.75, 1.25, 2.25):
            H = blqh_amd.Hamiltonian(B, blqh_amd.params(NPHI, d=d, **PAR))
            p = H.lanczos(lam=300, seed=1, check_every=1)
            w = H.lanczos(lam=300, start=net.start_vector(B, d, PAR["delta_sas"]), check_every=1)
            assert p["converged"] == 1 and w["converged"] == 1
            assert w["E0"] == pytest.approx(p["E0"], rel=1e-10, abs=0)
            r = H.spmv(w["psi"]) - w["E0"] * w["psi"]
            assert np.linalg.norm(r) < 1e-6
            plain_it.append(p["iterations"])
            warm_it.append(w["iterations"])
    print("iterations plain", plain_it, "warm", warm_it, flush=True)
    assert np.mean(warm_it) <= 0.95 * np.mean(plain_it)
