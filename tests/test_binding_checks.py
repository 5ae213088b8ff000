"""Argument checks of the Python binding that run before any library call:
a start vector, state or input of the wrong length must raise ValueError
instead of letting the C side read past the caller's buffer (the C entry
points take no length for these arrays, include/blqh.h)."""
import numpy as np
import pytest

import blqh_amd


class _Fake:
    def __init__(self, n):
        self.n = n


def _ham(n):
    h = object.__new__(blqh_amd.Hamiltonian)
    h.n = n
    return h


def test_lanczos_start_length():
    with pytest.raises(ValueError, match="start vector has 5 entries"):
        _ham(6).lanczos(start=np.ones(5, np.complex128))


def test_reconstruct_lengths():
    h = _ham(6)
    with pytest.raises(ValueError, match="start vector"):
        h.reconstruct(np.ones((1, 3)), start=np.ones(7))
    with pytest.raises(ValueError, match="alpha"):
        h.reconstruct(np.ones((1, 3)), alpha=np.ones(1))
    with pytest.raises(ValueError, match="beta"):
        h.reconstruct(np.ones((1, 3)), beta=np.ones(2))
    with pytest.raises(ValueError, match="coefficient vectors"):
        h.reconstruct(np.ones((17, 3)))


def test_spmv_length():
    with pytest.raises(ValueError, match="x has 3 entries"):
        _ham(4).spmv(np.ones(3))


def test_batch_start_lengths():
    b = object.__new__(blqh_amd.Batch)
    b.hams = [_Fake(4), _Fake(5)]
    with pytest.raises(ValueError, match="1 start vectors for 2 problems"):
        b.lanczos(starts=[np.ones(4)])
    with pytest.raises(ValueError, match="start vector 1 has 4 entries"):
        b.lanczos(starts=[None, np.ones(4)])
    with pytest.raises(ValueError, match="3 seeds for 2 problems"):
        b.lanczos(seeds=[1, 2, 3])


def test_observables_and_mlp_lengths():
    with pytest.raises(ValueError, match="state has 2 entries"):
        blqh_amd.observables(_Fake(3), np.ones(2))
    m = object.__new__(blqh_amd.Mlp)
    m.nparams = 10
    with pytest.raises(ValueError, match="9 parameters"):
        m.params = np.ones(9)
    with pytest.raises(ValueError, match="one length"):
        m.loss_grad(np.ones(3, np.uint64), np.ones(3), np.ones(2), np.ones(3), np.ones(3))
    with pytest.raises(ValueError, match="state 0 has 2 entries"):
        m.train([(_Fake(3), 0.0, 0.0, np.ones(2))])


def test_record_lengths(tmp_path):
    with pytest.raises(ValueError, match="record arrays"):
        blqh_amd.record_save(str(tmp_path / "r"), 4, 1, np.ones(3), np.ones(3), np.ones(3), np.ones(3), np.ones(3))
