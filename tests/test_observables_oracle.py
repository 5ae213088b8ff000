"""The oracle's restatement of the reference's ground-state observables,
pinned by sum rules and by the Hellmann-Feynman theorem: the tunnelling part
of H is -Delta_SAS T, so <psi|T|psi> = -dE0/dDelta_SAS (finite differences
of dense ground-state energies)."""
import numpy as np
import pytest

import oracle as O


def _ground(sec, nphi, d, dsas, **kw):
    B = O.Basis(sec)
    H = O.Hamiltonian(B, O.params(nphi, d=d, dsas=dsas, **kw))
    w, v = np.linalg.eigh(H.dense())
    return B, w, v[:, 0]


@pytest.mark.parametrize("sec,nphi,d,dsas", [
    (dict(S=2, nphi=6, nel=6, nel_up=-1, J=3, K=3), 6, 1.0, 0.05),
    (dict(S=2, nphi=6, nel=6, nel_up=-1, J=1, K=2), 6, 0.5, 0.3),
    (dict(S=4, nphi=4, nel=4, nel_up=-1, nel_su=-1, J=0, K=1), 4, 1.0, 0.05),
])
def test_sum_rules_and_hellmann_feynman(sec, nphi, d, dsas):
    s = O.sector(**sec)
    h = 1e-4
    B, w, psi = _ground(s, nphi, d, dsas)
    ob = O.observables(B, psi)
    nel = sec["nel"]
    occ = ob["occupation"]
    if sec["S"] == 2:
        # every electron sits in one of the 2 nphi orbitals
        assert occ.sum() == pytest.approx(nel, rel=1e-12)
        assert ob["pseudospin_sz"] == pytest.approx((2 * occ[0].sum() - nel) / (2 * nel), abs=1e-12)
    ep = np.linalg.eigvalsh(O.Hamiltonian(O.Basis(s), O.params(nphi, d=d, dsas=dsas + h)).dense())[0]
    em = np.linalg.eigvalsh(O.Hamiltonian(O.Basis(s), O.params(nphi, d=d, dsas=dsas - h)).dense())[0]
    hf = abs((ep - em) / (2 * h)) / nel
    assert w[1] - w[0] > 1e-6  # non-degenerate ground state
    assert hf > 0.1
    assert ob["pseudospin_sx"] == pytest.approx(hf, rel=1e-6)


def test_fixed_layer_counts_have_no_tunnelling():
    """With the upper-layer count fixed every tunnelling target leaves the basis
    (the reference drops it), so Sx vanishes."""
    s = O.sector(S=4, nphi=4, nel=4, nel_up=2, nel_su=2, J=0, K=1)
    B, w, psi = _ground(s, 4, 1.0, 0.05)
    assert O.observables(B, psi)["pseudospin_sx"] == 0.0


def test_spinflip_is_half_of_sx_under_lower_layer_spin_flip_symmetry():
    """Exchanging spin up and down in the lower layer maps the same-spin hops
    onto the spin-flip hops and leaves H invariant when Delta_Z = 0 and
    theta_B = 0 (nphi even, so both tunnelling kinds carry the same
    empty-orbital sign convention).  A non-degenerate ground state of a sector
    with all spin counts is then an eigenvector of that exchange, so
    <T_sf> = <T_ss> and spinflip = Sx / 2."""
    s = O.sector(S=4, nphi=4, nel=4, nel_up=-1, nel_su=-1, J=0, K=1)
    B, w, psi = _ground(s, 4, 1.0, 0.05)
    assert w[1] - w[0] > 1e-6
    ob = O.observables(B, psi)
    assert ob["pseudospin_sx"] > 0.1
    assert ob["spinflip"] == pytest.approx(ob["pseudospin_sx"] / 2, rel=1e-10)


def test_spinflip_vanishes_when_spins_are_polarised():
    """With every electron spin-up (nel_su = nel) each spin-flip target leaves
    the basis, while the same-spin tunnelling still acts."""
    s = O.sector(S=4, nphi=4, nel=4, nel_up=-1, nel_su=4, J=2, K=2)
    B, w, psi = _ground(s, 4, 1.0, 0.05)
    ob = O.observables(B, psi)
    assert ob["spinflip"] == 0.0
    assert ob["pseudospin_sx"] > 0.3


def test_spinless_has_no_spinflip():
    s = O.sector(S=2, nphi=6, nel=6, nel_up=-1, J=3, K=3)
    B, w, psi = _ground(s, 6, 1.0, 0.05)
    assert O.observables(B, psi)["spinflip"] == 0.0
