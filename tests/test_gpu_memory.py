"""Device block reuse: blocks freed by destroyed handles are kept for
allocations of the same size and returned by blqh_release_cached."""
import numpy as np
import pytest

import blqh_amd

pytestmark = pytest.mark.gpu


def test_freed_blocks_are_reused_and_released():
    print("block cache", flush=True)
    B = blqh_amd.Basis(10, 10, -1, 1, 2)
    p = blqh_amd.params(10, d=1.0, delta_sas=0.05)
    blqh_amd.release_cached()
    H = blqh_amd.Hamiltonian(B, p)
    ref = H.export()
    e_ref = H.lanczos(lam=200, seed=4)["E0"]
    del H
    H2 = blqh_amd.Hamiltonian(B, p)  # same sizes: served from the kept blocks
    for a, b in zip(ref, H2.export()):
        assert np.array_equal(a, b)
    assert H2.lanczos(lam=200, seed=4)["E0"] == e_ref
    del H2
    assert blqh_amd.release_cached() > 0
    assert blqh_amd.release_cached() == 0
