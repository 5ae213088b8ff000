"""Lanczos spectra (tridiagonal eigenpairs, ghost filter) and the Lanczos
record file (the reference's save_to_file / read_from_file format).  Host
functions of the library: they run without a GPU, except the last test."""
import numpy as np
import pytest

import blqh_amd


def _tri(m, seed):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=m)
    b = np.concatenate([[1.0], np.abs(rng.normal(size=m - 1)) + 0.1, [0.0]])[:m + 1]
    return a, b


@pytest.mark.parametrize("m", [1, 2, 7, 60, 200])
def test_tridiagonal_eigenpairs_match_lapack(m):
    a, b = _tri(m, m)
    T = np.diag(a) + np.diag(b[1:m], 1) + np.diag(b[1:m], -1)
    w, V = np.linalg.eigh(T)
    eig, first, vec0 = blqh_amd.tridiag_spectrum(a, b)
    scale = max(1.0, np.abs(w).max())
    np.testing.assert_allclose(eig, w, rtol=0, atol=1e-12 * scale)
    np.testing.assert_allclose(np.abs(first), np.abs(V[0, :]), rtol=0, atol=1e-10)
    assert np.linalg.norm(T @ vec0 - eig[0] * vec0) < 1e-10 * scale
    assert vec0[0] >= 0 and np.linalg.norm(vec0) == pytest.approx(1.0, abs=1e-12)


def test_split_tridiagonal_and_repeated_eigenvalues():
    """A zero off-diagonal splits T into blocks whose spectra coincide: the QL
    deflation path and exact repeats (Lanczos ghosts look like this)."""
    a = np.array([1.0, 2.0, 3.0, 1.0, 2.0, 3.0])
    b = np.array([1.0, 0.5, 0.5, 0.0, 0.5, 0.5, 0.0])
    eig, first, vec0 = blqh_amd.tridiag_spectrum(a, b)
    T = np.diag(a) + np.diag(b[1:6], 1) + np.diag(b[1:6], -1)
    np.testing.assert_allclose(eig, np.linalg.eigvalsh(T), atol=1e-13)
    assert abs(eig[1] - eig[0]) < 1e-12  # the two blocks repeat each level


def test_ghost_filter_keeps_jumps_larger_than_gap():
    # consecutive differences: 1e-7 (ghost), 0.1, 2e-6 (ghost), 0.1, 1e-5 (not > gap), 0.05
    e = np.array([-1.0, -1.0 + 1e-7, -0.9 + 1e-7, -0.9 + 2.1e-6, -0.8 + 2.1e-6, -0.8 + 1.21e-5, -0.75])
    got = blqh_amd.spectrum_distinct(e, kmax=7, gap=1e-5)
    np.testing.assert_array_equal(got, [e[0], e[2], e[4], e[6]])
    np.testing.assert_array_equal(blqh_amd.spectrum_distinct(e, kmax=2, gap=1e-5), [e[0], e[2]])


def test_record_text_format(tmp_path):
    p = tmp_path / "rec.txt"
    blqh_amd.record_save(p, 79, 1, eig=[-5.09335459, -3.25], beta=[1.0, 0.5, 0.25], alpha=[-4.0, -2.125],
                         psi0=[0.9, 0.4358898943540674], psin0=[0.9, -0.4358898943540674])
    expected = ("nHilbert:=" + " " * 10 + "79\n"
                "lambda:=" + " " * 10 + "2\n"
                "seed:=" + " " * 10 + "1\n"
                "eigenvalues:=" + " " * 7 + " [ -5.0933546 , -3.25 ,  ] \n"
                "norm:=" + " " * 7 + " [ 1 , 0.5 , 0.25 ,  ] \n"
                "overlap:=" + " " * 7 + " [ -4 , -2.125 ,  ] \n"
                "psi_0:=" + " " * 7 + " [ 0.9 , 0.43588989 ,  ] \n"
                "psi_n0:=" + " " * 7 + " [ 0.9 , -0.43588989 ,  ] \n")
    assert p.read_text() == expected


def test_record_round_trip(tmp_path):
    a, b = _tri(40, 3)
    eig, first, vec0 = blqh_amd.tridiag_spectrum(a, b)
    p = tmp_path / "r.txt"
    blqh_amd.record_save(p, 123456, 77, eig, b, a, vec0, first)
    r = blqh_amd.record_load(p)
    assert (r["nhilbert"], r["lambda"], r["seed"]) == (123456, 40, 77)
    for key, v in (("eigenvalues", eig), ("norm", b), ("overlap", a), ("psi_0", vec0), ("psi_n0", first)):
        np.testing.assert_array_equal(r[key], [float(f"{x:.8g}") for x in v])
    with pytest.raises(blqh_amd.BlqhError):
        blqh_amd.record_load(p, cap=10)


@pytest.mark.gpu
def test_low_lying_spectrum_from_fixed_lambda_lanczos():
    """Fixed-lambda Lanczos (no reorthogonalisation) on the 79-state README
    sector: the lowest ghost-filtered Ritz level is E0 and every low-lying
    distinct level of the dense spectrum appears among the filtered levels
    (the 1e-5 filter of the reference can keep a not yet converged ghost
    between them, so the lists are not compared one to one)."""
    B = blqh_amd.Basis(8, 8, 4, 4, 4)
    H = blqh_amd.Hamiltonian(B, blqh_amd.params(8, d=3.0))
    r = H.lanczos(lam=200, seed=1, tol=0.0, want_psi=False)
    eig, _, _ = blqh_amd.tridiag_spectrum(r["alpha"], r["beta"])
    lev = blqh_amd.spectrum_distinct(eig, kmax=8, gap=1e-5)
    rp, col, val = H.export()
    import scipy.sparse as sp
    w = np.linalg.eigvalsh(sp.csr_matrix((val, col, rp), shape=(H.n, H.n)).toarray())
    ref = blqh_amd.spectrum_distinct(w, kmax=3, gap=1e-5)
    assert lev[0] == pytest.approx(w[0], rel=1e-12)
    for e in ref:
        assert np.min(np.abs(lev - e)) < 1e-8 * abs(e), e


def test_tridiag_vectors_match_lapack_and_reject_bad_ranges():
    a, b = _tri(30, 5)
    T = np.diag(a) + np.diag(b[1:30], 1) + np.diag(b[1:30], -1)
    w, V = np.linalg.eigh(T)
    vecs = blqh_amd.tridiag_vectors(a, b, 2, 4)
    for k in range(4):
        v = vecs[k]
        assert v[0] >= 0
        assert abs(abs(v @ V[:, 2 + k]) - 1) < 1e-10
        assert np.linalg.norm(T @ v - w[2 + k] * v) < 1e-10
    for k0, cnt in ((-1, 1), (29, 2), (0, 0), (0, 31)):
        with pytest.raises(blqh_amd.BlqhError):
            blqh_amd.tridiag_vectors(a, b, k0, cnt)


def test_spectrum_distinct_edge_cases():
    e = np.array([-1.0, -1.0, -0.5])
    assert len(blqh_amd.spectrum_distinct(e, kmax=0)) == 0
    np.testing.assert_array_equal(blqh_amd.spectrum_distinct(e[:1], kmax=5), [-1.0])
    np.testing.assert_array_equal(blqh_amd.spectrum_distinct(e, kmax=5, gap=0.6), [-1.0])


def test_record_single_step_and_malformed_files(tmp_path):
    p = tmp_path / "one.txt"
    blqh_amd.record_save(p, 1, 3, eig=[-2.5], beta=[1.0, 0.0], alpha=[-2.5], psi0=[1.0], psin0=[1.0])
    r = blqh_amd.record_load(p)
    assert r["lambda"] == 1 and r["nhilbert"] == 1 and r["seed"] == 3
    np.testing.assert_array_equal(r["norm"], [1.0, 0.0])
    text = p.read_text()
    for bad in ("", "nHilbert:= x\n", text.replace("[ -2.5 ,", "[ abc ,"), text.split("norm")[0]):
        q = tmp_path / "bad.txt"
        q.write_text(bad)
        with pytest.raises(blqh_amd.BlqhError):
            blqh_amd.record_load(q)
    with pytest.raises(blqh_amd.BlqhError):
        blqh_amd.record_load(tmp_path / "missing.txt")
