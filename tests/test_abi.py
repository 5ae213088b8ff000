"""The C ABI: libblqh.so loads and exports every entry point include/blqh.h
declares; host-only entry points work without a GPU."""
import os
import re

import numpy as np
import pytest

import blqh_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "blqh.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(blqh_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_declared_symbols():
    L = blqh_amd.lib()
    names = declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
    assert set(names) <= set(blqh_amd._SIGS)


def test_start_vector_matches_oracle(oracle_mod):
    # 100003 entries: 200006 draws, 321 state blocks of the generator
    for n, seed in ((1, 0), (79, 1), (1000, 4242), (100003, 5)):
        np.testing.assert_array_equal(blqh_amd.start_vector(n, seed), oracle_mod.start_vector(n, seed))


def test_tridiag_lowest_vs_lapack():
    rng = np.random.default_rng(0)
    for m in (1, 2, 3, 10, 80, 300):
        a = rng.normal(size=m)
        b = rng.uniform(0.1, 1.0, size=m - 1)
        e, v = blqh_amd.tridiag_lowest(a, b)
        T = np.diag(a) + np.diag(b, 1) + np.diag(b, -1)
        w, V = np.linalg.eigh(T)
        assert e == pytest.approx(w[0], abs=1e-12 * max(1, abs(w[0])))
        assert abs(abs(v @ V[:, 0]) - 1) < 1e-10


def test_invalid_arguments_return_error_codes():
    with pytest.raises(blqh_amd.BlqhError) as e:
        blqh_amd.tridiag_lowest([], [])
    assert e.value.code == -1


def test_null_batch_arguments_are_refused():
    """The batch entry points check their pointers before touching a device."""
    lib = blqh_amd.lib()
    res = (blqh_amd.LanczosResult * 1)()
    opts = blqh_amd.LanczosOpts(200, 0, 1e-10, 1, 1, None, 1e-7, 0, 0.0)
    assert lib.blqh_batch_lanczos_continue(None, blqh_amd.C.byref(opts), None, 0.0, res, None) == -1
    assert lib.blqh_batch_lanczos(None, blqh_amd.C.byref(opts), None, None, res, None) == -1
