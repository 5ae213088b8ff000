"""The benchmark workload at its full size (spinless bilayer N_phi = 16,
N_e = 16, all layer densities, (J, K) = (0, 0), d = 1.0, Delta_SAS = 0.02:
2,348,125 states), where the oracle cannot follow, checked through the
properties that do not depend on size: the sector dimension, Hermiticity
and linearity of the SpMV on random vectors, Lanczos ground states from two
seeds that agree (energy and state), and the eigen-residual of the state
(a real sector, so psi = sum c_l conj(v_l) is an eigenvector of H)."""
import numpy as np
import pytest

import blqh_amd

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sector16():
    print("building the 2,348,125-state sector", flush=True)
    B = blqh_amd.Basis(16, 16, -1, 0, 0)
    H = blqh_amd.Hamiltonian(B, blqh_amd.params(16, d=1.0, delta_sas=0.02))
    print(f"n={H.n} nnz={H.nnz}", flush=True)
    return B, H


def test_dimensions(sector16):
    B, H = sector16
    assert H.n == 2348125
    # every row holds its diagonal; a few hundred entries per row on average
    assert 300 * H.n < H.nnz < 400 * H.n


def test_spmv_hermitian_and_linear(sector16):
    B, H = sector16
    rng = np.random.default_rng(7)
    x = rng.normal(size=H.n) + 1j * rng.normal(size=H.n)
    y = rng.normal(size=H.n) + 1j * rng.normal(size=H.n)
    Hx, Hy = H.spmv(x), H.spmv(y)
    a, b = np.vdot(x, Hy), np.vdot(Hx, y)
    assert abs(a - b) <= 1e-12 * abs(a)
    z = H.spmv(2.5 * x - 1j * y)
    ref = 2.5 * Hx - 1j * Hy
    assert np.linalg.norm(z - ref) <= 1e-13 * np.linalg.norm(ref)


def test_ground_state_from_two_seeds(sector16):
    B, H = sector16
    r1 = H.lanczos(lam=200, seed=1, tol=1e-10)
    r2 = H.lanczos(lam=200, seed=2, tol=1e-10)
    print({k: r1[k] for k in ("E0", "iterations", "residual")}, flush=True)
    assert r1["converged"] == 1 and r2["converged"] == 1
    assert r2["E0"] == pytest.approx(r1["E0"], rel=1e-12)
    assert abs(np.vdot(r1["psi"], r2["psi"])) >= 1 - 1e-10
    psi = r1["psi"]
    res = np.linalg.norm(H.spmv(psi) - r1["E0"] * psi)
    assert res < 1e-6
    assert r1["E_rq"] == pytest.approx(r1["E0"], rel=1e-12)
