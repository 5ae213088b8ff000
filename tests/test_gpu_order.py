"""The SpMV formats' internal row order changes no result: the same sector
built with the order on (default) and off (BLQH_ORDER=0) exports the same CSR,
gives the same SpMV (reference-order vectors in and out) to rounding, and
the same ground state.  Multi-panel real sector and single-panel complex
sector."""
import os

import numpy as np
import pytest

import blqh_amd

pytestmark = pytest.mark.gpu

CASES = [((14, 14, -1, 0, 0), dict(d=1.0, delta_sas=0.05)),   # 204,660 states, 4 column panels, real
         ((12, 12, -1, 1, 2), dict(d=0.6, delta_sas=0.1))]    # 18,806 states, complex


def _pair(case, kw):
    nphi, nel, nu, J, K = case
    par = blqh_amd.params(nphi, **kw)
    Bo = blqh_amd.Basis(nphi, nel, nu, J, K)
    Ho = blqh_amd.Hamiltonian(Bo, par)
    old = os.environ.get("BLQH_ORDER")
    os.environ["BLQH_ORDER"] = "0"
    try:
        Bi = blqh_amd.Basis(nphi, nel, nu, J, K)
        Hi = blqh_amd.Hamiltonian(Bi, par)
    finally:
        if old is None:
            del os.environ["BLQH_ORDER"]
        else:
            os.environ["BLQH_ORDER"] = old
    return Ho, Hi


@pytest.mark.parametrize("case,kw", CASES, ids=["multipanel_real", "complex"])
def test_internal_order_changes_nothing(case, kw):
    print(f"case {case}", flush=True)
    Ho, Hi = _pair(case, kw)
    ro, co, vo = Ho.export()
    ri, ci, vi = Hi.export()
    assert np.array_equal(ro, ri) and np.array_equal(co, ci) and np.array_equal(vo, vi)
    rng = np.random.default_rng(3)
    x = rng.normal(size=Ho.n) + 1j * rng.normal(size=Ho.n)
    yo, yi = Ho.spmv(x), Hi.spmv(x)
    scale = np.abs(vo).max() * np.abs(x).max() * np.diff(ro).max()
    np.testing.assert_allclose(yo, yi, rtol=0, atol=1e-14 * scale)
    a = Ho.lanczos(lam=300, seed=5)
    b = Hi.lanczos(lam=300, seed=5)
    assert a["converged"] == 1 and b["converged"] == 1
    assert a["E0"] == pytest.approx(b["E0"], rel=1e-10, abs=0)
    assert a["E_rq"] == pytest.approx(np.vdot(a["psi"], Ho.spmv(a["psi"])).real, rel=1e-12)
    # psi = sum c_l conj(v_l) (the reference's convention): conj(psi) is H's eigenvector
    g = np.conj(a["psi"])
    r = Ho.spmv(g) - a["E0"] * g
    assert np.linalg.norm(r) < 1e-5 * abs(a["E0"])
    # replay (three passes) in the internal order as well
    c = Ho.lanczos(lam=300, seed=5, store_basis=0)
    assert c["passes"] == 3 and c["E0"] == a["E0"]
    assert abs(np.vdot(a["psi"], c["psi"])) >= 1 - 1e-12
