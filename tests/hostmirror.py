"""ctypes access to tests/hostcheck/_build/libhostcheck.so, the CPU mirror of
the device pipeline (same per-element code as the kernels)."""
import ctypes as C
import os

import numpy as np

import blqh_amd

PATH = os.path.join(os.path.dirname(__file__), "hostcheck", "_build", "libhostcheck.so")
_L = None


def lib():
    global _L
    if _L is None:
        _L = C.CDLL(PATH)
        _L.hc_basis.restype = C.c_int64
        _L.hc_basis.argtypes = [C.POINTER(blqh_amd.Sector), C.c_void_p, C.c_void_p, C.c_int64]
        _L.hc_hamil.restype = C.c_int64
        _L.hc_hamil.argtypes = [C.POINTER(blqh_amd.Sector), C.POINTER(blqh_amd.Params), C.c_void_p, C.c_void_p,
                                C.c_void_p, C.c_int64]
        _L.hc_hamil_ql_major.restype = C.c_int64
        _L.hc_hamil_ql_major.argtypes = _L.hc_hamil.argtypes
        _L.hc_sector_words.restype = C.c_int64
        _L.hc_sector_words.argtypes = [C.POINTER(blqh_amd.Sector), C.c_int, C.c_void_p, C.c_int64]
        _L.hc_internal_order.restype = None
        _L.hc_internal_order.argtypes = [C.POINTER(blqh_amd.Sector), C.c_void_p, C.c_int64, C.c_void_p]
        _L.hc_hash_order_check.restype = C.c_int
        _L.hc_hash_order_check.argtypes = [C.POINTER(blqh_amd.Sector), C.POINTER(C.c_int64)]
        _L.hc_tridiag_vector_k.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
        _L.hc_excited_mix.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        _L.hc_hash_order_levels_check.restype = C.c_int
        _L.hc_hash_order_levels_check.argtypes = [C.POINTER(blqh_amd.Sector), C.POINTER(C.c_int64)]
        _L.hc_mod_by_rcp.restype = None
        _L.hc_mod_by_rcp.argtypes = [C.c_int, C.c_int, C.c_float, C.c_void_p]
        _L.hc_xcd_lists.restype = None
        _L.hc_xcd_lists.argtypes = [C.c_int] + [C.c_void_p] * 6
    return _L


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def basis(sec):
    cap = 1 << 20
    ids = np.empty(cap, np.uint64)
    per = np.empty(cap, np.int16)
    n = lib().hc_basis(C.byref(sec), _p(ids), _p(per), cap)
    assert n >= 0
    return ids[:n].copy(), per[:n].copy()


def hamil(sec, par, n, cap):
    rp = np.empty(n + 1, np.int64)
    col = np.empty(cap, np.int64)
    val = np.empty(2 * cap)
    nnz = lib().hc_hamil(C.byref(sec), C.byref(par), _p(rp), _p(col), _p(val), cap)
    assert nnz >= 0, "capacity too small"
    return rp, col[:nnz].copy(), val[:2 * nnz].view(np.complex128).copy()


def hamil_ql_major(sec, par, n, cap):
    """The mirror's Hamiltonian in the reference's loop order (translation,
    then pair and shift) instead of the build kernel's grouping."""
    rp = np.empty(n + 1, np.int64)
    col = np.empty(cap, np.int64)
    val = np.empty(2 * cap)
    nnz = lib().hc_hamil_ql_major(C.byref(sec), C.byref(par), _p(rp), _p(col), _p(val), cap)
    assert nnz >= 0, "capacity too small"
    return rp, col[:nnz].copy(), val[:2 * nnz].view(np.complex128).copy()


def sector_words(sec, with_spin=True):
    cap = 1 << 22
    out = np.empty(cap, np.uint64)
    m = lib().hc_sector_words(C.byref(sec), 1 if with_spin else 0, _p(out), cap)
    assert m >= 0, "rank/unrank inconsistency"
    return out[:m].copy()


def xcd_lists(item0, nnz, npan, real):
    """The batched K1's per-XCD work lists: (items, seg) as the library builds them."""
    item0 = np.ascontiguousarray(item0, np.int64)
    nnz = np.ascontiguousarray(nnz, np.int64)
    npan = np.ascontiguousarray(npan, np.int32)
    real = np.ascontiguousarray(real, np.int32)
    items = np.empty(2 * int(item0[-1]), np.int32)
    seg = np.empty(27, np.int32)
    lib().hc_xcd_lists(len(nnz), _p(item0), _p(nnz), _p(npan), _p(real), _p(items), _p(seg))
    return items, seg


def mod_by_rcp(n, d, r):
    """rowterms.hpp mod_by_rcp(s, d, r) for s = 0 .. n-1."""
    out = np.empty(n, np.int32)
    lib().hc_mod_by_rcp(n, d, float(np.float32(r)), _p(out))
    return out


def internal_order(sec, ids):
    """panel.hip internal_order: internal position of every row (None: identity)."""
    ids = np.ascontiguousarray(ids, np.uint64)
    out = np.empty(len(ids), np.int32)
    lib().hc_internal_order(C.byref(sec), _p(ids), len(ids), _p(out))
    return None if len(ids) and out[0] < 0 else out


def hash_order_check(sec):
    """(replayed hash-map order == real std::unordered_map order, words)."""
    n = C.c_int64()
    ok = lib().hc_hash_order_check(C.byref(sec), C.byref(n))
    return bool(ok), n.value


def hash_order_levels_check(sec):
    """(level-by-level order, the device algorithm == real map order, words)."""
    n = C.c_int64()
    ok = lib().hc_hash_order_levels_check(C.byref(sec), C.byref(n))
    return bool(ok), n.value


def tridiag_vector_k(a, b, k):
    """Eigenvector k (ascending) of the symmetric tridiagonal (a, b), as the batched solver computes it."""
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.empty(a.size)
    lib().hc_tridiag_vector_k(a.size, a.ctypes.data, b.ctypes.data, k, x.ctypes.data)
    return x


def excited_mix(a, b, c0):
    """The batched solver's excited mix of a tridiagonal with ground vector c0."""
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    c0 = np.ascontiguousarray(c0, np.float64)
    out = np.empty(a.size)
    lib().hc_excited_mix(a.size, a.ctypes.data, b.ctypes.data, c0.ctypes.data, out.ctypes.data)
    return out
