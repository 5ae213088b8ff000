"""Batched Lanczos (blqh_batch_*): problems of different sizes, real and
complex, spinless and spinful, one and several column panels, solved together
must give what the single-problem solver gives for each of them (energy to
1e-10 relative, state overlap >= 1 - 1e-10), and the batch itself must be
deterministic (same bits for any replay length, run to run, any lane count)."""
import numpy as np
import pytest
import scipy.sparse as sp

import blqh_amd

pytestmark = pytest.mark.gpu

# (name, nspecies, nphi, nel, nel_up, nel_su, J, K, d, delta_sas)
PROBLEMS = [
    ("c1_real", 2, 8, 8, 4, -1, 4, 4, 3.0, 0.0),
    ("all_density_complex", 2, 8, 8, -1, -1, 1, 3, 1.0, 0.05),
    ("u5_real", 2, 10, 10, 5, -1, 0, 0, 0.5, 0.0),
    ("u5_complex", 2, 10, 10, 5, -1, 3, 2, 2.0, 0.0),
    ("spinful_complex", 4, 6, 6, 3, 3, 0, 1, 1.0, 0.05),
    ("four_panels_real", 2, 14, 14, -1, -1, 0, 0, 1.0, 0.05),
]


@pytest.fixture(scope="module")
def hams():
    out = []
    for name, S, nphi, nel, nu, nsu, J, K, d, dsas in PROBLEMS:
        print(f"build {name}", flush=True)
        B = blqh_amd.Basis(nphi, nel, nu, J, K, nspecies=S, nel_su=nsu)
        H = blqh_amd.Hamiltonian(B, blqh_amd.params(nphi, d=d, delta_sas=dsas))
        out.append((name, B, H))
    return out


@pytest.fixture(scope="module")
def batch(hams):
    return blqh_amd.Batch([H for _, _, H in hams])


@pytest.fixture(scope="module")
def batch_result(batch):
    return batch.lanczos(lam=250, seeds=[11 + i for i in range(len(PROBLEMS))], check_every=4)


def test_batch_matches_single_solver(hams, batch_result):
    for i, (name, B, H) in enumerate(hams):
        print(f"compare {name} n={H.n}", flush=True)
        s = H.lanczos(lam=250, seed=11 + i, check_every=4)
        b = batch_result[i]
        assert b["converged"] == 1 and s["converged"] == 1, name
        assert b["E0"] == pytest.approx(s["E0"], rel=1e-10, abs=0), name
        ov = abs(np.vdot(b["psi"], s["psi"]))
        assert ov >= 1 - 1e-10, (name, ov)
        assert np.linalg.norm(b["psi"]) == pytest.approx(1.0, abs=1e-12)
        assert b["E_rq"] == pytest.approx(np.vdot(b["psi"], H.spmv(b["psi"])).real, rel=1e-12, abs=1e-12)
        assert abs(b["iterations"] - s["iterations"]) <= 2, name


def test_batch_matches_dense_diagonalisation(hams, batch_result):
    """Small problems: the batched ground state is the lowest eigenpair of the
    exported CSR (LAPACK through numpy)."""
    for i, (name, B, H) in enumerate(hams):
        if H.n > 4000:
            continue
        rp, col, val = H.export()
        A = sp.csr_matrix((val, col, rp), shape=(H.n, H.n)).toarray()
        w, v = np.linalg.eigh(A)
        b = batch_result[i]
        assert b["E0"] == pytest.approx(w[0], rel=1e-10, abs=0), name
        # psi = sum c_l conj(v_l): the ground state of conj(H) = H^T, i.e. conj of H's
        r = np.linalg.norm(A.conj() @ b["psi"] - w[0] * b["psi"])
        assert r < 1e-6, (name, r)


def test_replay_length_does_not_change_bits(batch, batch_result):
    seeds = [11 + i for i in range(len(PROBLEMS))]
    for chk in (1, 7):
        print(f"check_every {chk}", flush=True)
        r = batch.lanczos(lam=250, seeds=seeds, check_every=chk)
        for a, b in zip(r, batch_result):
            assert a["iterations"] == b["iterations"]
            assert a["E0"] == b["E0"]
            assert np.array_equal(a["psi"], b["psi"])


def test_batch_after_member_relayout(hams, batch, batch_result):
    """Pinning a member to the jagged SpMV reorders its panels into new
    buffers; the batch must pick up the current ones on its next solve."""
    name, B, H = hams[-1]
    H.set_spmv(3, 64, 512, 2, 2)
    H.spmv(np.ones(H.n, np.complex128))
    r = batch.lanczos(lam=250, seeds=[11 + i for i in range(len(PROBLEMS))], check_every=4, want_psi=False)
    for a, b in zip(r, batch_result):
        assert a["E0"] == b["E0"]
    H.set_spmv(-1)


def test_lane_counts_agree(batch, batch_result):
    seeds = [11 + i for i in range(len(PROBLEMS))]
    keep = batch.lanes
    try:
        for lanes in (4, 32, 64):
            batch.lanes = lanes
            r = batch.lanczos(lam=250, seeds=seeds, check_every=4, want_psi=False)
            for a, b in zip(r, batch_result):
                assert a["E0"] == pytest.approx(b["E0"], rel=1e-11, abs=0)
    finally:
        batch.lanes = keep
    with pytest.raises(blqh_amd.BlqhError):
        batch.lanes = 12
    with pytest.raises(blqh_amd.BlqhError):
        blqh_amd.Batch(batch.hams, lanes=12)


def test_batch_created_with_fixed_lanes(hams, batch_result):
    b = blqh_amd.Batch([H for _, _, H in hams], lanes=16)
    assert b.lanes == 16
    r = b.lanczos(lam=250, seeds=[11 + i for i in range(len(PROBLEMS))], check_every=4, want_psi=False)
    for a, c in zip(r, batch_result):
        assert a["E0"] == pytest.approx(c["E0"], rel=1e-11, abs=0)
    # a second solve with another lambda re-allocates the pools
    r2 = b.lanczos(lam=260, seeds=[11 + i for i in range(len(PROBLEMS))], check_every=4, want_psi=False)
    for a, c in zip(r2, r):
        assert a["E0"] == c["E0"] and a["iterations"] == c["iterations"]


def test_fixed_lambda_runs_exactly_lambda_steps(hams, batch):
    r = batch.lanczos(lam=40, seeds=3, tol=0.0, check_every=6, want_psi=False)
    for i, (name, B, H) in enumerate(hams):
        s = H.lanczos(lam=40, seed=3, tol=0.0, want_psi=False)
        assert r[i]["iterations"] == s["iterations"], name  # 40 unless an invariant subspace ends it
        assert r[i]["E0"] == pytest.approx(s["E0"], rel=1e-10, abs=0), name


def test_warm_start_from_ground_state(hams, batch, batch_result):
    """Starting from the converged states (conjugated back to Krylov space:
    psi = sum c_l conj(v_l)) converges in a few steps to the same energies."""
    starts = [np.conj(b["psi"]) for b in batch_result]
    r = batch.lanczos(lam=250, seeds=0, starts=starts, check_every=2)
    for a, b in zip(r, batch_result):
        assert a["converged"] == 1
        assert a["iterations"] < b["iterations"]
        assert a["E0"] == pytest.approx(b["E0"], rel=1e-10, abs=0)


def test_single_problem_batch(hams):
    name, B, H = hams[2]
    r = blqh_amd.Batch([H]).lanczos(lam=200, seeds=5)
    s = H.lanczos(lam=200, seed=5)
    assert r[0]["E0"] == pytest.approx(s["E0"], rel=1e-10, abs=0)


def test_rebind_to_next_parameter_point(hams, batch):
    """A sweep's next point: same sectors, new d; the rebound batch gives what
    a freshly created batch gives, and a size mismatch is refused."""
    new = []
    for (name, S, nphi, nel, nu, nsu, J, K, d, dsas), (_, B, _) in zip(PROBLEMS, hams):
        new.append(blqh_amd.Hamiltonian(B, blqh_amd.params(nphi, d=d + 0.25, delta_sas=dsas)))
    seeds = [11 + i for i in range(len(PROBLEMS))]
    fresh = blqh_amd.Batch(new, lanes=batch.lanes).lanczos(lam=250, seeds=seeds, check_every=4, want_psi=False)
    old_members = batch.hams
    try:
        batch.rebind(new)
        r = batch.lanczos(lam=250, seeds=seeds, check_every=4, want_psi=False)
        for a, b in zip(r, fresh):
            assert a["E0"] == b["E0"] and a["iterations"] == b["iterations"]
        with pytest.raises(blqh_amd.BlqhError):
            batch.rebind(new[::-1])
    finally:
        batch.rebind(old_members)


def test_many_members_one_sector_each():
    """All 100 (J, K) sectors of the N_phi = 10 all-density system in one
    batch (1,846-1,855 states each): enough members that each is worked on
    whole, real and complex ones mixed.  Every member's ground state equals the
    single solver's; a second batch of the same members gives the same bits."""
    hs = []
    for J in range(10):
        for K in range(10):
            B = blqh_amd.Basis(10, 10, -1, J, K)
            hs.append(blqh_amd.Hamiltonian(B, blqh_amd.params(10, d=1.0, delta_sas=0.05)))
    print(f"batch of {len(hs)} sectors", flush=True)
    seeds = [3 + i for i in range(len(hs))]
    r = blqh_amd.Batch(hs, lanes=64).lanczos(lam=200, seeds=seeds, want_psi=False)
    r2 = blqh_amd.Batch(hs, lanes=64).lanczos(lam=200, seeds=seeds, want_psi=False)
    for i, H in enumerate(hs):
        s = H.lanczos(lam=200, seed=seeds[i])
        assert r[i]["converged"] == 1 and s["converged"] == 1, i
        assert r[i]["E0"] == pytest.approx(s["E0"], rel=1e-10, abs=0), i
        assert r2[i]["E0"] == r[i]["E0"] and r2[i]["iterations"] == r[i]["iterations"], i


@pytest.mark.parametrize("mix", [0.0, 0.05])
def test_continuation_matches_host_starts(hams, mix):
    """blqh_batch_lanczos_continue after a rebind starts every problem from
    conj(psi) of the previous solve kept on the device; it must agree with
    the same solve started from the host copies (starts = conj(psi), same
    start_mix and seeds), and it is refused before any solve and for another
    lambda."""
    seeds = [11 + i for i in range(len(PROBLEMS))]
    seeds2 = [40 + i for i in range(len(PROBLEMS))]
    bt = blqh_amd.Batch([H for _, _, H in hams])
    with pytest.raises(blqh_amd.BlqhError):
        bt.lanczos(lam=250, seeds=seeds, check_every=4, continue_prev=True)
    cold = bt.lanczos(lam=250, seeds=seeds, check_every=4)
    new = []
    for (name, S, nphi, nel, nu, nsu, J, K, d, dsas), (_, B, _) in zip(PROBLEMS, hams):
        new.append(blqh_amd.Hamiltonian(B, blqh_amd.params(nphi, d=d + 0.25, delta_sas=dsas)))
    bt.rebind(new)
    with pytest.raises(blqh_amd.BlqhError):
        bt.lanczos(lam=200, seeds=seeds2, check_every=4, continue_prev=True)
    cont = bt.lanczos(lam=250, seeds=seeds2, check_every=4, start_mix=mix, continue_prev=True)
    host = blqh_amd.Batch(new, lanes=bt.lanes).lanczos(
        lam=250, seeds=seeds2, check_every=4, starts=[np.conj(c["psi"]) for c in cold], start_mix=mix)
    for (name, _, H), a, b in zip(hams, cont, host):
        print(f"continuation {name} mix={mix}: it {a['iterations']} host {b['iterations']}", flush=True)
        assert a["converged"] == 1 and b["converged"] == 1, name
        assert a["E0"] == pytest.approx(b["E0"], rel=1e-10, abs=0), name
        assert abs(a["iterations"] - b["iterations"]) <= 4, name
        assert abs(np.vdot(a["psi"], b["psi"])) >= 1 - 1e-10, name
    # and again from the continuation's own states (no rebind): a few steps
    again = bt.lanczos(lam=250, seeds=seeds2, check_every=2, start_mix=mix, continue_prev=True, want_psi=False)
    for a, b in zip(again, cont):
        assert a["converged"] == 1
        assert a["E0"] == pytest.approx(b["E0"], rel=1e-10, abs=0)
    bt.close()


def test_continuation_with_excited_mix(hams):
    """excited_mix adds the previous solve's second and third Ritz vectors to
    the continuation start: the solve still ends at the cold ground state of
    the new point, and repeats bit for bit."""
    seeds = [11 + i for i in range(len(PROBLEMS))]
    bt = blqh_amd.Batch([H for _, _, H in hams])
    bt.lanczos(lam=250, seeds=seeds, check_every=4, want_psi=False)
    new = []
    for (name, S, nphi, nel, nu, nsu, J, K, d, dsas), (_, B, _) in zip(PROBLEMS, hams):
        new.append(blqh_amd.Hamiltonian(B, blqh_amd.params(nphi, d=d + 0.5, delta_sas=dsas)))
    cold = blqh_amd.Batch(new, lanes=bt.lanes).lanczos(lam=250, seeds=seeds, check_every=4, want_psi=False)
    bt.rebind(new)
    with pytest.raises(blqh_amd.BlqhError):
        bt.lanczos(lam=250, seeds=seeds, continue_prev=True, excited_mix=-1.0)
    bt2 = blqh_amd.Batch([H for _, _, H in hams], lanes=bt.lanes)
    bt2.lanczos(lam=250, seeds=seeds, check_every=4, want_psi=False)
    bt2.rebind(new)
    r1 = bt.lanczos(lam=250, seeds=seeds, check_every=4, start_mix=0.1, continue_prev=True, excited_mix=0.5)
    r2 = bt2.lanczos(lam=250, seeds=seeds, check_every=4, start_mix=0.1, continue_prev=True, excited_mix=0.5)
    for (name, _, H), a, b, c in zip(hams, r1, r2, cold):
        print(f"excited mix {name}: it {a['iterations']} cold {c['iterations']}", flush=True)
        assert a["converged"] == 1, name
        assert a["E0"] == pytest.approx(c["E0"], rel=1e-10, abs=0), name
        assert a["E0"] == b["E0"] and a["iterations"] == b["iterations"] and np.array_equal(a["psi"], b["psi"]), name
    bt.close()
    bt2.close()
