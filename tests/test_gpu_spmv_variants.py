"""Every SpMV variant the autotuner can pick, on multi-panel sectors (more
than 65536 states), against scipy's CSR product of the exported matrix (the
CSR itself is pinned to the oracle elsewhere)."""
import numpy as np
import pytest
import scipy.sparse as sp

import blqh_amd

pytestmark = pytest.mark.gpu

PLANS = [(0, l, 256, 0, 0) for l in (4, 8, 16, 32, 64)] + \
        [(1, l, t, w, u) for l in (4, 8, 16) for t in (512, 1024) for w in (1, 2) for u in (2, 4)] + \
        [(2, 64, t, w, u) for t in (512, 1024) for w in (1, 2, 4) for u in (1, 2, 4, 8)] + \
        [(3, 64, t, w, u) for t in (256, 512, 1024) for w in (1, 2, 4) for u in (1, 2, 4, 8)] + \
        [(4, 64, 1024, 1, 0)] + \
        [(5, R, t, w, u) for R, u in ((2, 1), (2, 2), (2, 4), (4, 1), (4, 2), (8, 1)) for t in (512, 1024)
         for w in (1, 2)]

SECTORS = [
    ("real_4panels", (2, 14, 14, -1, -1, 0, 0), dict(d=1.0, delta_sas=0.05)),
    ("complex_4panels", (2, 14, 14, -1, -1, 0, 3), dict(d=0.5, delta_sas=0.05)),
    ("halfK_stored_real", (2, 14, 14, 7, -1, 7, 7), dict(d=1.5)),
]


@pytest.fixture(scope="module", params=SECTORS, ids=[s[0] for s in SECTORS])
def built(request):
    name, (S, nphi, nel, nu, nsu, J, K), kw = request.param
    print(f"sector {name}", flush=True)
    B = blqh_amd.Basis(nphi, nel, nu, J, K, nspecies=S, nel_su=nsu)
    H = blqh_amd.Hamiltonian(B, blqh_amd.params(nphi, **kw))
    rp, col, val = H.export()
    A = sp.csr_matrix((val, col, rp), shape=(H.n, H.n))
    return name, B, H, A


@pytest.mark.parametrize("plan", PLANS, ids=[f"k{p[0]}_l{p[1]}_t{p[2]}_w{p[3]}_u{p[4]}" for p in PLANS])
def test_variant_matches_csr_product(built, plan):
    name, B, H, A = built
    print(f"{name} plan {plan}", flush=True)
    H.set_spmv(*plan)
    assert H.get_spmv()[:2] == plan[:2]
    rng = np.random.default_rng(1)
    x = rng.normal(size=H.n) + 1j * rng.normal(size=H.n)
    y = H.spmv(x)
    ref = A @ x
    scale = np.abs(A).max() * np.abs(x).max() * np.diff(A.indptr).max()
    np.testing.assert_allclose(y, ref, rtol=0, atol=2e-15 * scale)


def test_layout_round_trip(built):
    """Kind 3 reads the panels in step-major order, kind 5 in step-major
    order with R rows per lane, kinds 1-2 in row-major order; switching back
    and forth reorders the entries on the device and must give bit-identical
    products each time, whatever the order of the switches."""
    name, B, H, A = built
    print(f"{name} layout round trip", flush=True)
    rng = np.random.default_rng(2)
    x = rng.normal(size=H.n) + 1j * rng.normal(size=H.n)
    H.set_spmv(3, 64, 512, 2, 2)
    y3 = H.spmv(x)
    H.set_spmv(2, 64, 512, 2, 2)
    y2 = H.spmv(x)
    H.set_spmv(5, 4, 512, 2, 2)
    y5 = H.spmv(x)
    H.set_spmv(3, 64, 512, 2, 2)
    assert np.array_equal(H.spmv(x), y3)
    H.set_spmv(5, 4, 512, 2, 2)
    assert np.array_equal(H.spmv(x), y5)
    H.set_spmv(5, 8, 512, 2, 1)
    H.spmv(x)
    H.set_spmv(2, 64, 512, 2, 2)
    assert np.array_equal(H.spmv(x), y2)
    H.set_spmv(5, 4, 512, 2, 2)
    assert np.array_equal(H.spmv(x), y5)
    rp, col, val = H.export()
    assert np.array_equal(rp, A.indptr) and np.array_equal(col, A.indices) and np.array_equal(val, A.data)


def test_solver_independent_of_variant(built):
    name, B, H, A = built
    print(f"{name} solver variants", flush=True)
    H.set_spmv(0, 16)
    e_csr = H.lanczos(lam=300, seed=4, want_psi=False)["E0"]
    H.set_spmv(1, 8, 1024, 1, 4)
    e_pan = H.lanczos(lam=300, seed=4, want_psi=False)["E0"]
    H.set_spmv(-1)
    e_auto = H.lanczos(lam=300, seed=4, want_psi=False)["E0"]
    assert e_pan == pytest.approx(e_csr, rel=1e-11)
    H.set_spmv(4, 64, 1024, 1, 0)
    e_t4 = H.lanczos(lam=300, seed=4, want_psi=False)["E0"]
    assert e_t4 == pytest.approx(e_csr, rel=1e-11)
    H.set_spmv(5, 8, 1024, 1, 1)
    e_rows = H.lanczos(lam=300, seed=4, want_psi=False)["E0"]
    assert e_rows == pytest.approx(e_csr, rel=1e-11)
    H.set_spmv(3, 64, 512, 2, 2)
    e_jag = H.lanczos(lam=300, seed=4, want_psi=False)["E0"]
    assert e_auto == pytest.approx(e_csr, rel=1e-11)
    assert e_jag == pytest.approx(e_csr, rel=1e-11)
    assert H.get_spmv()[0] == 3


@pytest.mark.parametrize("R,u", [(2, 1), (4, 2), (8, 1)])
def test_rows_per_lane_on_small_matrices(R, u):
    """Kind 5 when a 64R-row block extends past the last row (79 and 1846
    states, real and complex)."""
    for args, kw in (((8, 8, 4, 4, 4), dict(d=3.0)), ((10, 10, 5, 3, 2), dict(d=2.0))):
        B = blqh_amd.Basis(*args)
        H = blqh_amd.Hamiltonian(B, blqh_amd.params(args[0], **kw))
        print(f"n={H.n} R={R}", flush=True)
        rp, col, val = H.export()
        A = sp.csr_matrix((val, col, rp), shape=(H.n, H.n))
        H.set_spmv(5, R, 512, 1, u)
        rng = np.random.default_rng(3)
        x = rng.normal(size=H.n) + 1j * rng.normal(size=H.n)
        ref = A @ x
        scale = np.abs(A).max() * np.abs(x).max() * np.diff(A.indptr).max()
        np.testing.assert_allclose(H.spmv(x), ref, rtol=0, atol=2e-15 * scale)
        e = H.lanczos(lam=300, seed=1)["E0"]
        H.set_spmv(0, 16)
        assert e == pytest.approx(H.lanczos(lam=300, seed=1)["E0"], rel=1e-11)


def test_greedy_step_order_on_request(built, monkeypatch):
    """BLQH_JAGGED_ORDER=greedy puts a single matrix's step-major panels in the
    greedy step order (at each step the lanes are served line by line, the x
    cache line wanted by most first); the plain step order is the default.
    Both give the CSR product, each order reproduces its own products bit for
    bit after a switch away and back, and the CSR is untouched."""
    name, B, H, A = built
    print(f"{name} greedy step order", flush=True)
    rng = np.random.default_rng(3)
    x = rng.normal(size=H.n) + 1j * rng.normal(size=H.n)
    ref = A @ x
    scale = np.abs(A).max() * np.abs(x).max() * np.diff(A.indptr).max()
    monkeypatch.setenv("BLQH_JAGGED_ORDER", "greedy")
    H.set_spmv(2, 64, 512, 2, 2)
    H.set_spmv(3, 64, 512, 2, 2)
    yg = H.spmv(x)
    monkeypatch.delenv("BLQH_JAGGED_ORDER")
    H.set_spmv(2, 64, 512, 2, 2)
    H.set_spmv(3, 64, 512, 2, 2)
    yp = H.spmv(x)
    np.testing.assert_allclose(yg, ref, rtol=0, atol=2e-15 * scale)
    np.testing.assert_allclose(yp, ref, rtol=0, atol=2e-15 * scale)
    monkeypatch.setenv("BLQH_JAGGED_ORDER", "greedy")
    H.set_spmv(2, 64, 512, 2, 2)
    H.set_spmv(3, 64, 512, 2, 2)
    assert np.array_equal(H.spmv(x), yg)
    monkeypatch.delenv("BLQH_JAGGED_ORDER")
    H.set_spmv(5, 4, 512, 2, 2)
    H.set_spmv(3, 64, 512, 2, 2)
    assert np.array_equal(H.spmv(x), yp)
    rp, col, val = H.export()
    assert np.array_equal(rp, A.indptr) and np.array_equal(col, A.indices) and np.array_equal(val, A.data)
