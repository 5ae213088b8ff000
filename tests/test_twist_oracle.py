"""Twisted boundary conditions (theta_x in the interlayer Coulomb table,
theta_y phases on every momentum transfer, theta_B phases on the spinful
same-spin tunnelling; BLQH_without_spin/lanczos_hamiltonian.cpp:45-85,
peer_set_hamil; BLQH_with_spin/lanczos_hamiltonian.cpp:45-84, 718-837) in the
oracle, pinned by the symmetries any correct restatement has: the twisted
Hamiltonian stays Hermitian (a phase attached to a hop and not to its reverse
breaks it), it reduces to the untwisted one at zero twist, and in sectors with
real translation phases (K = 0) reversing theta_y and theta_B conjugates it
entry by entry (a twist entering with the wrong sign in one term breaks
that).  theta_x enters only the interlayer Coulomb table, as
cos((2 pi s + 2 theta_x) q / nphi) (spinful: (2 pi s - theta_x)), so shifting
it by pi (spinful: -2 pi) moves every entry to the next s; the reference then
reads the table at |s| (lanczos_hamiltonian.cpp:263), which is why its
spectrum is not even in theta_x."""
import numpy as np
import pytest

import oracle as O

SECTORS = [
    ("spinless", dict(S=2, nphi=6, nel=6, nel_up=-1, J=0, K=0), dict(d=1.0, dsas=0.1)),
    ("spinless_u", dict(S=2, nphi=8, nel=8, nel_up=4, J=4, K=0), dict(d=0.5)),
    ("spinful", dict(S=4, nphi=4, nel=4, nel_up=-1, nel_su=-1, J=1, K=0), dict(d=1.5, dsas=0.1, dz=0.01)),
]
TWISTS = [dict(theta_y=0.7), dict(theta_b=0.4), dict(theta_y=-1.1, theta_b=0.25)]


def _H(sec, kw, tw):
    B = O.Basis(O.sector(**sec))
    return O.Hamiltonian(B, O.params(sec["nphi"], **kw, **tw)).dense()


@pytest.mark.parametrize("name,sec,kw", SECTORS, ids=[s[0] for s in SECTORS])
def test_twisted_hamiltonian_symmetries(name, sec, kw):
    H0 = _H(sec, kw, {})
    assert np.array_equal(_H(sec, kw, dict(theta_x=0.0, theta_y=0.0, theta_b=0.0)), H0)
    scale = np.abs(H0).max()
    for tw in TWISTS:
        if sec["S"] == 2 and set(tw) == {"theta_b"}:
            continue  # theta_B acts on the spinful tunnelling only
        H = _H(sec, kw, tw)
        assert np.abs(H - H.conj().T).max() < 1e-13 * scale, tw
        assert np.abs(H - H0).max() > 1e-6 * scale, tw  # the twist does act
        Hm = _H(sec, kw, {k: -v for k, v in tw.items()})
        assert np.abs(Hm - H.conj()).max() < 1e-13 * scale, tw
        w, wm = np.linalg.eigvalsh(H), np.linalg.eigvalsh(Hm)
        np.testing.assert_allclose(w, wm, rtol=0, atol=1e-12 * scale)
    # theta_x: Hermitian and acting, through the interlayer table only
    H = _H(sec, kw, dict(theta_x=0.3, theta_y=0.2))
    assert np.abs(H - H.conj().T).max() < 1e-13 * scale
    assert np.abs(H - _H(sec, kw, dict(theta_y=0.2))).max() > 1e-6 * scale


@pytest.mark.parametrize("S,nphi,d", [(2, 6, 1.0), (2, 8, 0.3), (4, 4, 1.5)])
def test_theta_x_shifts_the_interlayer_table(S, nphi, d):
    sec = O.sector(S=S, nphi=nphi, nel=nphi, J=0, K=0)
    th = 0.37
    shift = np.pi if S == 2 else -2 * np.pi
    C0, _, _ = O.tables(sec, O.params(nphi, d=d))
    C, _, _ = O.tables(sec, O.params(nphi, d=d, theta_x=th))
    Cs, _, _ = O.tables(sec, O.params(nphi, d=d, theta_x=th + shift))
    np.testing.assert_array_equal(C[0], C0[0])  # intralayer table untouched
    assert np.abs(C[1] - C0[1]).max() > 1e-4
    np.testing.assert_allclose(Cs[1], np.roll(C[1], -1, axis=0), rtol=0, atol=1e-12 * np.abs(C[1]).max())


def test_theta_x_breaks_hermiticity_outside_K0():
    """The reference's theta_x twist (interlayer table read at |s|,
    BLQH_without_spin/lanczos_hamiltonian.cpp:263,284; spinful
    BLQH_with_spin/lanczos_hamiltonian.cpp:501-709): a hop and
    its reverse read the table at different separations, so H is Hermitian at
    K = 0 only (the translation sums symmetrise it there).  The oracle keeps
    this as the reference has it; the engine's CSR follows the oracle."""
    for S, nphi, K, herm in ((2, 8, 0, True), (2, 8, 1, False), (2, 8, -1, False), (4, 4, 1, False)):
        sec = dict(S=S, nphi=nphi, nel=nphi, nel_up=-1, J=0, K=K)
        if S == 4:
            sec["nel_su"] = -1
        Hd = _H(sec, dict(d=1.0, dsas=0.1), dict(theta_x=0.3))
        err = np.abs(Hd - Hd.conj().T).max() / np.abs(Hd).max()
        assert (err < 1e-13) == herm, (S, K, err)
        if not herm:
            assert err > 1e-5
