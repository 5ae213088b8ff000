"""Write tests/golden/basis_sha256.json: SHA-256 of the basis (id[] as
little-endian u64, then basis_C[] as i16) of the large sectors, computed on
the CPU by the test mirror (tests/hostcheck: the map order replayed with a
linked list, pinned to a real std::unordered_map by the CPU tests, and the
reference's survivor rule).  The GPU tests compare the device basis with it.

    python tests/golden/make_basis_sha.py      # needs the built hostcheck library
"""
import ctypes as C
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import blqh_amd  # noqa: E402
import hostmirror  # noqa: E402

SECTORS = [(2, 16, 16, -1, -1, 0, 0), (2, 16, 16, 8, -1, 0, 0)]


def main():
    out = {}
    for case in SECTORS:
        cap = 3_000_000
        ids = np.empty(cap, np.uint64)
        per = np.empty(cap, np.int16)
        n = hostmirror.lib().hc_basis(C.byref(blqh_amd.Sector(*case)), ids.ctypes.data_as(C.c_void_p),
                                      per.ctypes.data_as(C.c_void_p), cap)
        assert n > 0, (case, n)
        h = hashlib.sha256()
        h.update(ids[:n].astype("<u8").tobytes())
        h.update(per[:n].astype("<i2").tobytes())
        out["_".join(map(str, case))] = {"n": int(n), "sha256": h.hexdigest()}
        print(case, n, h.hexdigest(), flush=True)
    with open(os.path.join(HERE, "basis_sha256.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
