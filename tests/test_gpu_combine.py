"""H(d, Delta_SAS) = H(d, 0) + Delta_SAS T for parameter sweeps: the merge of a
tunnelling-free build and a tunnelling-only build equals the full build of
the reference's set_hamil at that Delta_SAS (same CSR structure, Coulomb and
diagonal values bit for bit, tunnelling values to rounding) and gives the
same ground state."""
import numpy as np
import pytest

import blqh_amd

pytestmark = pytest.mark.gpu

# (name, basis args, basis kwargs, params kwargs without delta_sas)
CASES = [
    ("spinless_complex", (10, 10, -1, 3, 2), {}, dict(d=1.0)),
    ("spinless_real", (10, 10, -1, 0, 0), {}, dict(d=0.4, delta_v=0.02)),
    ("spinful_thetaB", (6, 6, 3, 0, 1), dict(nspecies=4, nel_su=-1), dict(d=1.5, theta_b=0.3, delta_z=0.01)),
    ("fixed_layers", (8, 8, 4, 4, 4), {}, dict(d=3.0)),
]


@pytest.mark.parametrize("name,args,bkw,pkw", CASES, ids=[c[0] for c in CASES])
def test_combine_equals_full_build(name, args, bkw, pkw):
    print(name, flush=True)
    B = blqh_amd.Basis(*args, **bkw)
    nphi = args[0]
    base = blqh_amd.Hamiltonian(B, blqh_amd.params(nphi, **pkw), terms=1)
    tun = blqh_amd.Hamiltonian(B, blqh_amd.params(nphi, **pkw), terms=2)
    if name == "fixed_layers":
        assert tun.nnz == 0  # every tunnelling target leaves a fixed-layer basis
    rpb, colb, valb = base.export()
    for ds in (0.0, 0.05, 0.3):
        full = blqh_amd.Hamiltonian(B, blqh_amd.params(nphi, delta_sas=ds, **pkw))
        comb = blqh_amd.Hamiltonian.combine(base, tun, ds)
        rp, col, val = full.export()
        rp2, col2, val2 = comb.export()
        assert np.array_equal(rp, rp2) and np.array_equal(col, col2), ds
        scale = np.abs(val).max()
        np.testing.assert_allclose(val2, val, rtol=1e-14, atol=1e-15 * scale)
        # entries of the base are copied as they are
        rows = np.repeat(np.arange(full.n), np.diff(rp))
        key = rows.astype(np.int64) * full.n + col
        keyb = np.repeat(np.arange(full.n), np.diff(rpb)).astype(np.int64) * full.n + colb
        pos = np.searchsorted(key, keyb)
        assert np.array_equal(key[pos], keyb) and np.array_equal(val2[pos], valb)
        e_full = full.lanczos(lam=300, seed=2)["E0"]
        e_comb = comb.lanczos(lam=300, seed=2)["E0"]
        assert e_comb == pytest.approx(e_full, rel=1e-12), ds


def test_combine_rejects_mismatched_inputs():
    B = blqh_amd.Basis(8, 8, -1, 0, 0)
    base = blqh_amd.Hamiltonian(B, blqh_amd.params(8, d=1.0), terms=1)
    tun = blqh_amd.Hamiltonian(B, blqh_amd.params(8, d=1.0), terms=2)
    with_t = blqh_amd.Hamiltonian(B, blqh_amd.params(8, d=1.0, delta_sas=0.1))
    for a, b, ds in ((with_t, tun, 0.1), (base, base, 0.1), (base, tun, -0.1), (base, tun, float("nan"))):
        with pytest.raises(blqh_amd.BlqhError):
            blqh_amd.Hamiltonian.combine(a, b, ds)
    other = blqh_amd.Hamiltonian(B, blqh_amd.params(8, d=2.0, theta_b=0.2), terms=2)
    base2 = blqh_amd.Hamiltonian(B, blqh_amd.params(8, d=1.0, theta_b=0.2), terms=1)
    with pytest.raises(blqh_amd.BlqhError):  # theta_B phases depend on d
        blqh_amd.Hamiltonian.combine(base2, other, 0.1)
    B2 = blqh_amd.Basis(8, 8, -1, 1, 0)
    tun2 = blqh_amd.Hamiltonian(B2, blqh_amd.params(8, d=1.0), terms=2)
    with pytest.raises(blqh_amd.BlqhError):
        blqh_amd.Hamiltonian.combine(base, tun2, 0.1)
    # a Delta_SAS = 0 full build is a valid base; d may differ when theta_B = 0
    base0 = blqh_amd.Hamiltonian(B, blqh_amd.params(8, d=2.5))
    tun0 = blqh_amd.Hamiltonian(B, blqh_amd.params(8, d=1.0), terms=2)
    full = blqh_amd.Hamiltonian(B, blqh_amd.params(8, d=2.5, delta_sas=0.1))
    comb = blqh_amd.Hamiltonian.combine(base0, tun0, 0.1)
    assert np.array_equal(full.export()[1], comb.export()[1])


def test_parts_build_their_spmv_formats_on_first_use():
    """A terms-1 part (used by a merge, often never solved) builds its SpMV
    formats on its first product, or on prepare(): the same products either
    way, equal to the full build at Delta_SAS = 0, and prepare() twice is a
    no-op."""
    B = blqh_amd.Basis(10, 10, -1, 3, 2)
    par = blqh_amd.params(10, d=1.0)
    full = blqh_amd.Hamiltonian(B, par)
    lazy = blqh_amd.Hamiltonian(B, par, terms=1)
    ready = blqh_amd.Hamiltonian(B, par, terms=1).prepare().prepare()
    rng = np.random.default_rng(5)
    x = rng.normal(size=B.n) + 1j * rng.normal(size=B.n)
    y_lazy = lazy.spmv(x)  # first use: SpMV formats, then the autotune
    plan = lazy.get_spmv()
    ready.set_spmv(*plan)
    full.set_spmv(*plan)
    y_ready = ready.spmv(x)
    y_full = full.spmv(x)
    assert np.array_equal(y_lazy, y_ready)
    np.testing.assert_allclose(y_lazy, y_full, rtol=0, atol=1e-13 * np.abs(y_full).max())
    r1 = lazy.lanczos(lam=200, seed=3, tol=1e-10, want_psi=False)
    r2 = full.lanczos(lam=200, seed=3, tol=1e-10, want_psi=False)
    assert r1["E0"] == pytest.approx(r2["E0"], rel=1e-12)
