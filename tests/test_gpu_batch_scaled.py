"""Shared-scale batches (blqh_batch_create_scaled): the Delta_SAS points of
one d (BLQH_without_spin/main.cpp:133-165) as one group per sector whose
problems H0 + s_g T share one pass over the sector's entries.  Every problem
must give what the single solver gives on blqh_hamil_combine(H0, T, s_g)
(energy to 1e-10 relative, state overlap >= 1 - 1e-10 on non-degenerate
levels, E_rq the Rayleigh quotient of the returned state), also after a
rebind to the next d, with real and complex, spinless and spinful sectors
mixed; the shared pass is deterministic."""
import numpy as np
import pytest

import blqh_amd

pytestmark = pytest.mark.gpu

# (name, nspecies, nphi, nel, nel_up, nel_su, J, K)
SECTORS = [
    ("all_density_real", 2, 8, 8, -1, -1, 0, 0),
    ("all_density_complex", 2, 8, 8, -1, -1, 1, 3),
    ("nphi10_complex", 2, 10, 10, -1, -1, 3, 2),
    ("spinful_complex", 4, 4, 4, -1, -1, 1, 1),
    ("nphi12_real", 2, 12, 12, -1, -1, 0, 6),
]
SCALES = [0.0, 0.05, 0.1]


def _pair(sec, d, cache):
    name, S, nphi, nel, nu, nsu, J, K = sec
    if name not in cache:
        cache[name] = blqh_amd.Basis(nphi, nel, nu, J, K, nspecies=S, nel_su=nsu)
    B = cache[name]
    par = blqh_amd.params(nphi, d=d)
    return B, blqh_amd.Hamiltonian(B, par, terms=1), blqh_amd.Hamiltonian(B, par, terms=2)


@pytest.fixture(scope="module")
def bases():
    return {}


def _check(sectors, pairs, res, G, seeds, lam=300):
    for q, (sec, (B, H0, T)) in enumerate(zip(sectors, pairs)):
        for g, s in enumerate(SCALES[:G]):
            p = q * G + g
            H = blqh_amd.Hamiltonian.combine(H0, T, s)
            ref = H.lanczos(lam=lam, seed=seeds[p], check_every=4)
            b = res[p]
            print(f"{sec[0]} scale {s}: n={H.n} E0 {b['E0']:.12f} single {ref['E0']:.12f}", flush=True)
            assert b["converged"] == 1 and ref["converged"] == 1, (sec[0], s)
            assert b["E0"] == pytest.approx(ref["E0"], rel=1e-10, abs=0), (sec[0], s)
            assert b["E_rq"] == pytest.approx(np.vdot(b["psi"], H.spmv(b["psi"])).real, rel=1e-12, abs=1e-12)
            g_state = np.conj(b["psi"])  # psi = sum c_l conj(v_l)
            r = H.spmv(g_state) - b["E0"] * g_state
            assert np.linalg.norm(r) < 1e-5 * abs(b["E0"]), (sec[0], s)


def test_shared_scales_match_single_solver(bases):
    pairs = [_pair(sec, 1.0, bases) for sec in SECTORS]
    G = len(SCALES)
    bt = blqh_amd.Batch.scaled([p[1] for p in pairs], [p[2] for p in pairs], SCALES)
    assert len(bt.hams) == len(SECTORS) * G
    seeds = [20 + i for i in range(len(SECTORS) * G)]
    res = bt.lanczos(lam=300, seeds=seeds, check_every=3)
    _check(SECTORS, pairs, res, G, seeds)
    # deterministic: the same solve again, the same bits
    res2 = bt.lanczos(lam=300, seeds=seeds, check_every=3)
    for a, b in zip(res, res2):
        assert a["E0"] == b["E0"] and np.array_equal(a["psi"], b["psi"])
    # next d: rebind (the batch's own templates are rebuilt)
    pairs2 = [_pair(sec, 0.4, bases) for sec in SECTORS]
    bt.rebind_scaled([p[1] for p in pairs2], [p[2] for p in pairs2])
    res3 = bt.lanczos(lam=300, seeds=seeds, check_every=3)
    _check(SECTORS, pairs2, res3, G, seeds)
    bt.close()


def test_shared_scale_continuation(bases):
    """A continuation solve of a shared-scale batch after rebind_scaled
    agrees with the same solve from host start vectors conj(psi)."""
    secs = SECTORS[1:3]
    G = len(SCALES)
    pairs = [_pair(sec, 1.0, bases) for sec in secs]
    bt = blqh_amd.Batch.scaled([p[1] for p in pairs], [p[2] for p in pairs], SCALES)
    seeds = [5 + i for i in range(len(secs) * G)]
    cold = bt.lanczos(lam=300, seeds=seeds, check_every=3)
    pairs2 = [_pair(sec, 1.3, bases) for sec in secs]
    bt.rebind_scaled([p[1] for p in pairs2], [p[2] for p in pairs2])
    cont = bt.lanczos(lam=300, seeds=seeds, check_every=3, start_mix=0.05, continue_prev=True)
    host = bt.lanczos(lam=300, seeds=seeds, check_every=3, start_mix=0.05, starts=[np.conj(c["psi"]) for c in cold])
    for p, (a, b) in enumerate(zip(cont, host)):
        assert a["converged"] == 1 and b["converged"] == 1, p
        assert a["E0"] == pytest.approx(b["E0"], rel=1e-10, abs=0), p
        assert abs(np.vdot(a["psi"], b["psi"])) >= 1 - 1e-10, p
    _check(secs, pairs2, cont, G, seeds)
    bt.close()


@pytest.mark.parametrize("G", [1, 2, 4])
def test_group_sizes(bases, G):
    scales = [0.0, 0.05, 0.1, 0.2][:G]
    secs = SECTORS[1:3]
    pairs = [_pair(sec, 2.0, bases) for sec in secs]
    bt = blqh_amd.Batch.scaled([p[1] for p in pairs], [p[2] for p in pairs], scales)
    seeds = [5 + i for i in range(len(secs) * G)]
    res = bt.lanczos(lam=300, seeds=seeds, check_every=2)
    for q, (sec, (B, H0, T)) in enumerate(zip(secs, pairs)):
        for g, s in enumerate(scales):
            H = blqh_amd.Hamiltonian.combine(H0, T, s)
            ref = H.lanczos(lam=300, seed=seeds[q * G + g], check_every=4)
            assert res[q * G + g]["E0"] == pytest.approx(ref["E0"], rel=1e-10, abs=0), (sec[0], s)
    bt.close()


def test_shared_scale_refusals(bases):
    B, H0, T = _pair(SECTORS[0], 1.0, bases)
    with pytest.raises(ValueError):
        blqh_amd.Batch.scaled([H0], [T], [0.0, 0.1, 0.2, 0.3, 0.4])
    with pytest.raises(blqh_amd.BlqhError):
        blqh_amd.Batch.scaled([H0], [T], [-0.1])
    with pytest.raises(blqh_amd.BlqhError):  # tunnelling matrix in place of the base
        blqh_amd.Batch.scaled([T], [T], [0.1])
    big = blqh_amd.Basis(14, 14, -1, 0, 0)  # 204,660 states: more than one panel's 32768
    par = blqh_amd.params(14, d=1.0)
    with pytest.raises(blqh_amd.BlqhError):
        blqh_amd.Batch.scaled([blqh_amd.Hamiltonian(big, par, terms=1)], [blqh_amd.Hamiltonian(big, par, terms=2)],
                              [0.05])
    bt = blqh_amd.Batch.scaled([H0], [T], [0.05])
    with pytest.raises(blqh_amd.BlqhError):
        bt.rebind([H0])  # plain rebind of a shared-scale batch
    with pytest.raises(blqh_amd.BlqhError):
        bt.lanes = 8
    bt.close()


def test_config3_all_sectors_shared():
    """The sweep workload at its size: N_phi = 12, N_e = 12, all densities, all
    144 (J, K) sectors (18,772-18,806 states) at d = 0.6 with Delta_SAS in
    {0, 0.05, 0.1}: 432 problems in one shared-scale batch, each equal to the
    single solver on the merged Hamiltonian."""
    scales = [0.0, 0.05, 0.1]
    bases = [blqh_amd.Basis(12, 12, -1, J, K) for J in range(12) for K in range(12)]
    assert all(18772 <= B.n <= 18806 for B in bases)
    par = blqh_amd.params(12, d=0.6)
    base = [blqh_amd.Hamiltonian(B, par, terms=1) for B in bases]
    tun = [blqh_amd.Hamiltonian(B, par, terms=2) for B in bases]
    print("built 144 sectors", flush=True)
    bt = blqh_amd.Batch.scaled(base, tun, scales)
    seeds = [3 + i for i in range(432)]
    res = bt.lanczos(lam=250, seeds=seeds, check_every=2, want_psi=False)
    bt.close()
    print("shared-scale solve done", flush=True)
    for q in range(144):
        for g, s in enumerate(scales):
            p = 3 * q + g
            H = blqh_amd.Hamiltonian.combine(base[q], tun[q], s)
            ref = H.lanczos(lam=250, seed=seeds[p], check_every=4, want_psi=False)
            # a problem that needs more than 250 steps stops unconverged in both
            assert res[p]["converged"] == ref["converged"], (q, s)
            assert res[p]["E0"] == pytest.approx(ref["E0"], rel=1e-10, abs=0), (q, s)
            assert abs(res[p]["iterations"] - ref["iterations"]) <= 2, (q, s)
    assert sum(r["converged"] for r in res) >= 430
