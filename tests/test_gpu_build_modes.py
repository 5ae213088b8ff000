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


@pytest.mark.parametrize("name,args,kw,bkw", CASES, ids=[c[0] for c in CASES])
def test_row_hash_capacities_agree(name, args, kw, bkw):
    """A build over a basis whose earlier builds kept at most 192 entries per
    row hashes rows in 256 slots instead of 512; rows that do not fit go to
    the 4096-slot pass.  Forced 256 and 512 slots (BLQH_BUILD_CAP) and the
    default give the same CSR bit for bit; at N_phi = 14 (rows of up to
    ~260 entries) the forced 256-slot build sends rows through the overflow
    pass."""
    print(name, flush=True)
    B = blqh_amd.Basis(*args, **bkw)
    p = blqh_amd.params(args[0], **kw)
    first = blqh_amd.Hamiltonian(B, p).export()
    for cap in ("256", "512"):
        os.environ["BLQH_BUILD_CAP"] = cap
        try:
            again = blqh_amd.Hamiltonian(B, p).export()
        finally:
            del os.environ["BLQH_BUILD_CAP"]
        for a, b in zip(first, again):
            assert np.array_equal(a, b), cap
    again = blqh_amd.Hamiltonian(B, p).export()
    for a, b in zip(first, again):
        assert np.array_equal(a, b)
