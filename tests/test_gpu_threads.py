"""Library calls from two host threads at once: every host thread queues its
work on a stream of its own, so one thread's Hamiltonian builds overlap
another thread's batched solve, and released device blocks are handed to a
new owner only after the work queued before their release has finished.
Checked: matrices built on a worker thread while the main thread solves
equal (bit for bit) those built alone, and the concurrent batched solves give
the energies of the sequential ones."""
import threading
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import blqh_amd

pytestmark = pytest.mark.gpu

SECTORS = [(J, K) for J in range(0, 10, 3) for K in range(0, 10, 3)]


def _bases():
    return [blqh_amd.Basis(10, 10, -1, J, K) for J, K in SECTORS]


def _build(bases, d):
    return [blqh_amd.Hamiltonian(B, blqh_amd.params(10, d=d, delta_sas=0.05)) for B in bases]


def test_builds_on_a_worker_thread_overlap_solves():
    print("threads: bases", flush=True)
    bases = _bases()
    ds = [0.5, 1.0, 1.5, 2.0]
    ref_csr = {}
    ref_e = {}
    for d in ds:  # sequential reference: one thread
        hs = _build(bases, d)
        ref_csr[d] = [H.export() for H in hs]
        ref_e[d] = [r["E0"] for r in blqh_amd.Batch(hs).lanczos(seeds=list(range(1, len(hs) + 1)), want_psi=False)]
        del hs
    print("threads: pipelined", flush=True)
    main_id = threading.get_ident()
    ids = set()

    def work(d):
        ids.add(threading.get_ident())
        return _build(bases, d)

    with ThreadPoolExecutor(1) as ex:
        fut = ex.submit(work, ds[0])
        batch = None
        for k, d in enumerate(ds):
            hs = fut.result()
            if k + 1 < len(ds):
                fut = ex.submit(work, ds[k + 1])  # builds while this point solves
            for H, (rp, col, val) in zip(hs, ref_csr[d]):
                rp2, col2, val2 = H.export()
                np.testing.assert_array_equal(rp2, rp)
                np.testing.assert_array_equal(col2, col)
                np.testing.assert_array_equal(val2, val)
            if batch is None:
                batch = blqh_amd.Batch(hs)
            else:
                batch.rebind(hs)
            res = batch.lanczos(seeds=list(range(1, len(hs) + 1)), want_psi=False)
            assert all(r["converged"] == 1 for r in res)
            np.testing.assert_allclose([r["E0"] for r in res], ref_e[d], rtol=1e-12, atol=0)
    assert ids and main_id not in ids
