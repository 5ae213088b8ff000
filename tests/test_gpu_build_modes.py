"""The Hamiltonian build hashes every row once and stages the sorted rows
when the staging slots fit in HBM, else counts, scans and fills; both paths
must give the same CSR bit for bit (the two-pass path is forced with
BLQH_BUILD_TWO_PASS)."""
import os

import numpy as np
import pytest

import blqh_amd

pytestmark = pytest.mark.gpu

CASES = [
    ("readme", (8, 8, 4, 4, 4), dict(d=3.0), {}),
    ("all_density_complex", (10, 10, -1, 3, 2), dict(d=1.0, delta_sas=0.05), {}),
    ("spinful", (6, 6, 3, 0, 1), dict(d=1.0, delta_sas=0.05, delta_z=0.01), dict(nspecies=4, nel_su=3)),
    ("four_panels", (14, 14, -1, 0, 0), dict(d=1.0, delta_sas=0.05), {}),
]


@pytest.mark.parametrize("name,args,kw,bkw", CASES, ids=[c[0] for c in CASES])
def test_single_and_two_pass_builds_agree(name, args, kw, bkw):
    print(name, flush=True)
    B = blqh_amd.Basis(*args, **bkw)
    p = blqh_amd.params(args[0], **kw)
    one = blqh_amd.Hamiltonian(B, p).export()
    os.environ["BLQH_BUILD_TWO_PASS"] = "1"
    try:
        two = blqh_amd.Hamiltonian(B, p).export()
    finally:
        del os.environ["BLQH_BUILD_TWO_PASS"]
    for a, b in zip(one, two):
        assert np.array_equal(a, b)
