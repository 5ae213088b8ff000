"""The device basis, whose representative choice rests on the device's
level-by-level computation of the reference's std::unordered_map iteration
order (basis.hip map_positions), against the CPU mirror (tests/hostcheck),
which replays the map's node list on the host: id[] and basis_C[] equal
element by element on medium sectors, and by SHA-256 on the benchmark
sector (2,348,125 states from 37.6M words) and the fixed-layer N_phi = 16
sector, whose mirror bases take a minute on the CPU
(tests/golden/make_basis_sha.py)."""
import hashlib
import json
import os

import numpy as np
import pytest

import blqh_amd
from tests import hostmirror

pytestmark = pytest.mark.gpu

MEDIUM = [(2, 12, 12, -1, -1, 1, 2), (2, 12, 12, 6, -1, 0, 0), (2, 14, 14, 7, -1, 7, 7), (4, 8, 8, 4, 4, 3, 5),
          (4, 8, 8, 4, -1, 0, 0), (1, 18, 6, -1, -1, 3, 5), (1, 16, 8, -1, -1, 2, 7)]


@pytest.mark.skipif(not os.path.exists(hostmirror.PATH), reason="host mirror not built")
@pytest.mark.parametrize("case", MEDIUM, ids=lambda c: "_".join(map(str, c)))
def test_device_basis_equals_mirror(case):
    S, nphi, nel, nu, nsu, J, K = case
    print(f"basis {case}", flush=True)
    B = blqh_amd.Basis(nphi, nel, nu, J, K, nspecies=S, nel_su=nsu)
    ids, per = hostmirror.basis(blqh_amd.Sector(*case))
    assert B.n == len(ids)
    np.testing.assert_array_equal(B.id, ids)
    np.testing.assert_array_equal(B.period, per)


with open(os.path.join(os.path.dirname(__file__), "golden", "basis_sha256.json")) as _f:
    LARGE = json.load(_f)


@pytest.mark.parametrize("key", sorted(LARGE))
def test_large_basis_checksum(key):
    S, nphi, nel, nu, nsu, J, K = map(int, key.split("_"))
    print(f"basis {key}", flush=True)
    B = blqh_amd.Basis(nphi, nel, nu, J, K, nspecies=S, nel_su=nsu)
    assert B.n == LARGE[key]["n"]
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(B.id).astype("<u8").tobytes())
    h.update(np.ascontiguousarray(B.period).astype("<i2").tobytes())
    assert h.hexdigest() == LARGE[key]["sha256"]
