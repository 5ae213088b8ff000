"""blqh_lanczos_noalloc (no device allocation: the solver's static per-GPU
scratch) against blqh_lanczos with store_basis = 0, through the C ABI.

Two calls in a row on one GPU, the second on a smaller Hamiltonian, then an
untuned one (CSR kernel) and a caller-given start: energy to 1e-10 relative,
state overlap >= 1 - 1e-10, the same number of steps.  With the tuned plan in
place both entry points run the same kernels on the same start vector (the
seeded mt19937 one is drawn on the device here, on the host there), so the
Lanczos coefficients agree bit for bit."""
import numpy as np
import pytest

import blqh_amd

pytestmark = pytest.mark.gpu


def _compare(H, seed, start=None, exact=True):
    r = H.lanczos(lam=400, seed=seed, store_basis=0, start=start)
    q = H.lanczos_noalloc(lam=400, seed=seed, start=start)
    print(f"n={H.n} iterations {r['iterations']} / {q['iterations']}", flush=True)
    assert r["converged"] == 1 and q["converged"] == 1
    assert q["passes"] == 3
    assert q["iterations"] == r["iterations"]
    assert q["E0"] == pytest.approx(r["E0"], rel=1e-10, abs=0)
    assert q["E_rq"] == pytest.approx(r["E_rq"], rel=1e-10, abs=0)
    assert abs(np.vdot(r["psi"], q["psi"])) >= 1 - 1e-10
    if exact:
        np.testing.assert_array_equal(q["alpha"], r["alpha"])
        np.testing.assert_array_equal(q["beta"], r["beta"])
    return q


def test_noalloc_matches_lanczos_twice_then_smaller():
    # 18,806-state complex sector (N_phi = 12, all layer densities), then the
    # 79-state README sector: the second call reuses the same scratch
    B1 = blqh_amd.Basis(12, 12, -1, 1, 2)
    H1 = blqh_amd.Hamiltonian(B1, blqh_amd.params(12, d=1.0, delta_sas=0.05))
    q1 = _compare(H1, seed=7)
    B2 = blqh_amd.Basis(8, 8, 4, 4, 4)
    H2 = blqh_amd.Hamiltonian(B2, blqh_amd.params(8, d=3.0))
    assert H2.n < H1.n
    q2 = _compare(H2, seed=3)
    assert float("%.10g" % (q2["E_rq"] / 8)) == -1.235882219
    # the first Hamiltonian again after the smaller one: same result as before
    q3 = H1.lanczos_noalloc(lam=400, seed=7)
    np.testing.assert_array_equal(q3["alpha"], q1["alpha"])
    np.testing.assert_array_equal(q3["psi"], q1["psi"])


def test_noalloc_untuned_and_given_start():
    B = blqh_amd.Basis(10, 10, 5, 3, 2)
    H = blqh_amd.Hamiltonian(B, blqh_amd.params(10, d=2.0))
    # no solve has tuned this Hamiltonian yet: the CSR kernel runs
    q = H.lanczos_noalloc(lam=400, seed=11)
    r = H.lanczos(lam=400, seed=11, store_basis=0)
    assert q["converged"] == 1 and r["converged"] == 1
    assert q["E0"] == pytest.approx(r["E0"], rel=1e-10, abs=0)
    assert abs(np.vdot(r["psi"], q["psi"])) >= 1 - 1e-10
    rng = np.random.default_rng(2)
    start = rng.normal(size=H.n) + 1j * rng.normal(size=H.n)
    _compare(H, seed=0, start=start)


def test_noalloc_limits():
    B = blqh_amd.Basis(8, 8, 4, 4, 4)
    H = blqh_amd.Hamiltonian(B, blqh_amd.params(8, d=1.0))
    with pytest.raises(blqh_amd.BlqhError) as e:
        H.lanczos_noalloc(lam=1025)
    assert e.value.code == -4
    B = blqh_amd.Basis(14, 14, -1, 0, 0)
    assert B.n > 131072
    H = blqh_amd.Hamiltonian(B, blqh_amd.params(14, d=1.0))
    with pytest.raises(blqh_amd.BlqhError) as e:
        H.lanczos_noalloc(lam=10)
    assert e.value.code == -4
