"""Warm-start network on the GPU: forward/backward kernels against the
float64 oracle (oracle/mlp_oracle.py, itself pinned by finite differences),
inference against the oracle, and training on engine ground states that
lowers the loss and the Lanczos iteration count while every solve keeps the
plain solver's energy."""
import numpy as np
import pytest

import blqh_amd
from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


def _data(nbits, n, seed):
    rng = np.random.default_rng(seed)
    words = rng.integers(0, 1 << nbits, size=n, dtype=np.uint64)
    d = rng.uniform(0, 3, n).astype(np.float32)
    dsas = rng.choice([0.0, 0.05, 0.1], n).astype(np.float32)
    ya = rng.normal(-4, 2, n).astype(np.float32)
    ys = rng.integers(0, 2, n).astype(np.float32)
    return words, d, dsas, ya, ys


@pytest.mark.parametrize("nbits,n", [(28, 1000), (32, 333), (6, 64)])
def test_loss_and_gradient_match_oracle(nbits, n):
    print(f"nbits {nbits} samples {n}", flush=True)
    net = blqh_amd.Mlp(nbits, seed=3)
    rng = np.random.default_rng(nbits)
    theta = (net.params + rng.normal(0, 0.05, net.nparams)).astype(np.float32)
    net.params = theta
    assert np.array_equal(net.params, theta)
    data = _data(nbits, n, 5)
    L, g = net.loss_grad(*data)
    Lr, gr = M.loss_grad(theta.astype(np.float64), nbits, *[np.asarray(a, np.float64) if a.dtype != np.uint64 else a
                                                             for a in data])
    assert L == pytest.approx(Lr, rel=1e-5)
    # FP32 kernels against a float64 restatement: per-entry error bounded by the gradient's scale
    np.testing.assert_allclose(g, gr, rtol=0, atol=2e-5 * np.abs(gr).max())


def test_inference_matches_oracle():
    B = blqh_amd.Basis(8, 8, -1, 0, 0)
    net = blqh_amd.Mlp(16, seed=9)
    v = net.start_vector(B, 1.3, 0.05)
    r = M.start_vector(net.params.astype(np.float64), 16, B.id, 1.3, 0.05)
    assert np.all(v.imag == 0)
    assert abs(np.vdot(r, v.real)) >= 1 - 1e-6
    assert np.linalg.norm(v) == pytest.approx(1.0, abs=1e-12)


@pytest.fixture(scope="module")
def trained():
    """nphi = 10, Ne = 10, all layer densities, real sectors (0,0) and (5,5):
    ground states at d = 0.5 .. 2.5, Delta_SAS = 0.05, train the network."""
    nphi, nel, dsas = 10, 10, 0.05
    sectors = [(0, 0), (5, 5)]
    bases = {s: blqh_amd.Basis(nphi, nel, -1, *s) for s in sectors}
    sets = []
    for s in sectors:
        for d in (0.5, 1.0, 1.5, 2.0, 2.5):
            H = blqh_amd.Hamiltonian(bases[s], blqh_amd.params(nphi, d=d, delta_sas=dsas))
            r = H.lanczos(lam=300, seed=1, check_every=5)
            assert r["converged"] == 1
            sets.append((bases[s], d, dsas, r["psi"]))
    net = blqh_amd.Mlp(2 * nphi, seed=7)
    info = net.train(sets, steps=4000, batch=4096, lr=3e-3, lr_final=1e-4, seed=11)
    print("training", info, flush=True)
    return nphi, dsas, bases, net, info


def test_training_lowers_the_loss(trained):
    nphi, dsas, bases, net, info = trained
    assert info["steps"] == 4000
    assert info["loss_last"] < 0.3 * info["loss_first"]


def test_warm_start_saves_iterations_same_energy(trained):
    """At d between the training points: the network's start vector needs fewer
    Lanczos steps than the seeded random one, and both reach the same E0
    (the plain solve is the reference point)."""
    nphi, dsas, bases, net, info = trained
    plain_it, warm_it = [], []
    for s, B in bases.items():
        for d in (0.75, 1.25, 2.25):
            H = blqh_amd.Hamiltonian(B, blqh_amd.params(nphi, d=d, delta_sas=dsas))
            p = H.lanczos(lam=300, seed=1, check_every=1)
            w = H.lanczos(lam=300, start=net.start_vector(B, d, dsas), check_every=1)
            assert p["converged"] == 1 and w["converged"] == 1
            assert w["E0"] == pytest.approx(p["E0"], rel=1e-10, abs=0)
            # the warm state is a ground state (H real: H psi = E0 psi); overlap with the
            # plain state is not checked, some of these ground states are degenerate
            r = H.spmv(w["psi"]) - w["E0"] * w["psi"]
            assert np.linalg.norm(r) < 1e-6
            plain_it.append(p["iterations"])
            warm_it.append(w["iterations"])
    print("iterations plain", plain_it, "warm", warm_it, flush=True)
    # the start's overlap enters the step count only logarithmically: a modest,
    # consistent saving at every point
    assert np.mean(warm_it) <= 0.95 * np.mean(plain_it)
    assert sum(w <= p for w, p in zip(warm_it, plain_it)) >= len(plain_it) - 1


def test_start_overlap_with_ground_state(trained):
    """At a training point the network's start vector has a large overlap with
    the (gauge-fixed, real) ground state, far above a random vector's ~N^-1/2."""
    nphi, dsas, bases, net, info = trained
    B = bases[(5, 5)]
    H = blqh_amd.Hamiltonian(B, blqh_amd.params(nphi, d=2.0, delta_sas=dsas))
    r = H.lanczos(lam=300, seed=1, check_every=5)
    v = net.start_vector(B, 2.0, dsas)
    ov = abs(np.vdot(v, r["psi"]))
    print("overlap", ov, "random scale", 1 / np.sqrt(B.n), flush=True)
    assert ov > 0.5


def test_warm_start_in_a_batch(trained):
    nphi, dsas, bases, net, info = trained
    hs, starts = [], []
    for s, B in bases.items():
        hs.append(blqh_amd.Hamiltonian(B, blqh_amd.params(nphi, d=1.75, delta_sas=dsas)))
        starts.append(net.start_vector(B, 1.75, dsas))
    batch = blqh_amd.Batch(hs)
    warm = batch.lanczos(lam=300, seeds=1, starts=starts, check_every=2)
    plain = batch.lanczos(lam=300, seeds=1, check_every=2)
    for w, p in zip(warm, plain):
        assert w["E0"] == pytest.approx(p["E0"], rel=1e-10, abs=0)
        assert w["iterations"] < p["iterations"]
