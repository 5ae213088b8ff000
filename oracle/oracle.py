"""ctypes bindings for the oracle library (test infrastructure only)."""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")
REF_SPINLESS = os.path.join(HERE, "_ref", "libref_basis_spinless.so")
REF_SPINFUL = os.path.join(HERE, "_ref", "libref_basis_spinful.so")


class Sector(C.Structure):
    _fields_ = [("S", C.c_int), ("nphi", C.c_int), ("nel", C.c_int), ("nel_up", C.c_int),
                ("nel_su", C.c_int), ("J", C.c_int), ("K", C.c_int)]


class Params(C.Structure):
    _fields_ = [("lx", C.c_double), ("ly", C.c_double), ("nLL", C.c_int), ("d", C.c_double),
                ("dsas", C.c_double), ("dv", C.c_double), ("dz", C.c_double),
                ("theta_b", C.c_double), ("theta_x", C.c_double), ("theta_y", C.c_double)]


def geometry(nphi, gamma=1.0):
    """ly = sqrt(2 pi nphi / gamma), lx = gamma ly (BLQH_without_spin/main.cpp:20-21)."""
    ly = np.sqrt(nphi * 2.0 * np.pi / gamma)
    return ly * gamma, ly


def build(force=False):
    """Compile liboracle.so (and the reference basis builds when the tree exists)."""
    if force or not os.path.exists(LIB):
        subprocess.check_call(["make", "-s", "-C", HERE, "oracle"])
    if os.path.isdir("/root/reference"):
        subprocess.check_call(["make", "-s", "-C", HERE, "ref"])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        vp, i64, u64, dp = C.c_void_p, C.c_int64, C.c_uint64, C.POINTER(C.c_double)
        L.orc_basis_new.restype = vp
        L.orc_basis_new.argtypes = [C.POINTER(Sector)]
        L.orc_basis_adopt.restype = vp
        L.orc_basis_adopt.argtypes = [C.POINTER(Sector), vp, vp, i64]
        L.orc_basis_size.restype = i64
        L.orc_basis_size.argtypes = [vp]
        L.orc_basis_export.argtypes = [vp, vp, vp]
        L.orc_basis_find.restype = i64
        L.orc_basis_find.argtypes = [vp, u64]
        L.orc_basis_free.argtypes = [vp]
        L.orc_sector_configs.restype = i64
        L.orc_sector_configs.argtypes = [C.POINTER(Sector)]
        L.orc_sector_configs_export.argtypes = [C.POINTER(Sector), vp]
        L.orc_translate.restype = u64
        L.orc_translate.argtypes = [C.POINTER(Sector), u64, C.c_int, C.POINTER(C.c_int)]
        L.orc_inv_translate.restype = u64
        L.orc_inv_translate.argtypes = [C.POINTER(Sector), u64, C.c_int, C.POINTER(C.c_int)]
        L.orc_sign2.argtypes = [u64, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_sign1.argtypes = [u64, C.c_int, C.c_int]
        L.orc_tables.argtypes = [C.POINTER(Sector), C.POINTER(Params), vp, vp, vp]
        L.orc_hamil_new.restype = vp
        L.orc_hamil_new.argtypes = [vp, C.POINTER(Params), C.c_int]
        L.orc_hamil_nnz.restype = i64
        L.orc_hamil_nnz.argtypes = [vp]
        L.orc_hamil_n.restype = i64
        L.orc_hamil_n.argtypes = [vp]
        L.orc_hamil_export.argtypes = [vp, vp, vp, vp]
        L.orc_hamil_free.argtypes = [vp]
        L.orc_hamil_rows.restype = i64
        L.orc_hamil_rows.argtypes = [vp, C.POINTER(Params), vp, i64, C.c_int, vp, vp, vp, i64]
        L.orc_plant.argtypes = [C.c_int]
        L.orc_monolayer_new.restype = vp
        L.orc_monolayer_new.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
                                        C.POINTER(C.c_double)]
        L.orc_monolayer_basis.restype = i64
        L.orc_monolayer_basis.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, i64]
        L.orc_spmv32.argtypes = [i64, vp, vp, vp, vp, vp, C.c_int]
        L.orc_csr_wrap.restype = vp
        L.orc_csr_wrap.argtypes = [i64, vp, vp, vp]
        L.orc_spmv.argtypes = [vp, vp, vp]
        L.orc_start_vector.argtypes = [i64, C.c_uint, vp]
        L.orc_tridiag_eigh.argtypes = [C.c_int, vp, vp, vp, vp]
        L.orc_observables.restype = None
        L.orc_observables.argtypes = [vp, vp, vp]
        L.orc_lanczos.restype = C.c_double
        L.orc_lanczos.argtypes = [vp, C.c_int, C.c_uint, vp, vp, vp, vp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def sector(S=2, nphi=8, nel=8, nel_up=-1, nel_su=-1, J=-1, K=-1):
    return Sector(S, nphi, nel, nel_up, nel_su, J, K)


def params(nphi, gamma=1.0, nLL=0, d=0.0, dsas=0.0, dv=0.0, dz=0.0, theta_b=0.0, theta_x=0.0, theta_y=0.0):
    lx, ly = geometry(nphi, gamma)
    return Params(lx, ly, nLL, d, dsas, dv, dz, theta_b, theta_x, theta_y)


class Basis:
    """Sector basis built with the reference's rule (oracle restatement)."""

    def __init__(self, sec, adopt=None):
        """adopt=(ids, periods) skips the reduction and indexes the given list."""
        self.sec = sec
        if adopt is None:
            self.h = lib().orc_basis_new(C.byref(sec))
        else:
            ids = np.ascontiguousarray(adopt[0], np.uint64)
            per = np.ascontiguousarray(adopt[1], np.int16)
            self.h = lib().orc_basis_adopt(C.byref(sec), _p(ids), _p(per), len(ids))
        n = lib().orc_basis_size(self.h)
        self.id = np.empty(n, np.uint64)
        self.period = np.empty(n, np.int16)
        lib().orc_basis_export(self.h, _p(self.id), _p(self.period))

    def __len__(self):
        return len(self.id)

    def find(self, c):
        return lib().orc_basis_find(self.h, int(c))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_basis_free(self.h)
            self.h = None


def sector_configs(sec):
    n = lib().orc_sector_configs(C.byref(sec))
    out = np.empty(n, np.uint64)
    lib().orc_sector_configs_export(C.byref(sec), _p(out))
    return out


def translate(sec, c, k):
    s = C.c_int(0)
    r = lib().orc_translate(C.byref(sec), int(c), k, C.byref(s))
    return r, s.value


def inv_translate(sec, c, k):
    s = C.c_int(0)
    r = lib().orc_inv_translate(C.byref(sec), int(c), k, C.byref(s))
    return r, s.value


def sign2(c, n, m, nt, mt):
    return lib().orc_sign2(int(c), n, m, nt, mt)


def sign1(c, n, nt):
    return lib().orc_sign1(int(c), n, nt)


def tables(sec, par):
    nphi = sec.nphi
    Cm = np.empty(2 * nphi * nphi)
    ec = C.c_double(0)
    ft = np.empty(2 * nphi * nphi)
    lib().orc_tables(C.byref(sec), C.byref(par), _p(Cm), C.cast(C.byref(ec), C.c_void_p), _p(ft))
    return Cm.reshape(2, nphi, nphi), ec.value, ft.view(np.complex128).reshape(nphi, nphi)


class Hamiltonian:
    """Full CSR built row by row in the reference's accumulation order."""

    def __init__(self, basis, par, nthreads=0):
        self.basis = basis
        self.h = lib().orc_hamil_new(basis.h, C.byref(par), nthreads)
        self.n = lib().orc_hamil_n(self.h)
        nnz = lib().orc_hamil_nnz(self.h)
        self.rowptr = np.empty(self.n + 1, np.int64)
        self.col = np.empty(nnz, np.int64)
        self.val = np.empty(nnz, np.complex128)
        lib().orc_hamil_export(self.h, _p(self.rowptr), _p(self.col), _p(self.val))

    @property
    def nnz(self):
        return len(self.col)

    def dense(self):
        H = np.zeros((self.n, self.n), np.complex128)
        rows = np.repeat(np.arange(self.n), np.diff(self.rowptr))
        H[rows, self.col] = self.val
        return H

    def spmv(self, x):
        x = np.ascontiguousarray(x, np.complex128)
        y = np.empty_like(x)
        lib().orc_spmv(self.h, _p(x), _p(y))
        return y

    def lanczos(self, lam=200, seed=0):
        """Reference three-pass Lanczos, exactly lam steps.  Returns dict."""
        a = np.empty(lam)
        b = np.empty(lam + 1)
        e = np.empty(lam)
        psi = np.empty(self.n, np.complex128)
        E = lib().orc_lanczos(self.h, lam, seed, _p(a), _p(b), _p(e), _p(psi))
        return {"alpha": a, "beta": b, "eig": e, "psir": psi, "E_rq": E}

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_hamil_free(self.h)
            self.h = None


class CsrMatrix(Hamiltonian):
    """A caller's CSR (rowptr, col, val) with the oracle's SpMV and Lanczos."""

    def __init__(self, rowptr, col, val):
        self.basis = None
        self.rowptr = np.ascontiguousarray(rowptr, np.int64)
        self.col = np.ascontiguousarray(col, np.int64)
        self.val = np.ascontiguousarray(val, np.complex128)
        self.n = len(self.rowptr) - 1
        self.h = lib().orc_csr_wrap(self.n, _p(self.rowptr), _p(self.col), _p(self.val.view(np.float64)))


def hamil_rows(basis, par, rows, nthreads=0):
    """Build selected rows only.  Returns (rowptr, col, val)."""
    rows = np.ascontiguousarray(rows, np.int64)
    cap = max(1024, 1024 * len(rows))
    while True:
        rp = np.empty(len(rows) + 1, np.int64)
        col = np.empty(cap, np.int64)
        val = np.empty(2 * cap)
        tot = lib().orc_hamil_rows(basis.h, C.byref(par), _p(rows), len(rows), nthreads,
                                   _p(rp), _p(col), _p(val), cap)
        if tot >= 0:
            return rp, col[:tot], val[:2 * tot].view(np.complex128)
        cap = -tot


def start_vector(n, seed):
    out = np.empty(n, np.complex128)
    lib().orc_start_vector(n, seed, _p(out))
    return out


def tridiag_eigh(diag, off):
    m = len(diag)
    diag = np.ascontiguousarray(diag, np.float64)
    off = np.ascontiguousarray(off, np.float64)
    w = np.empty(m)
    V = np.empty((m, m))
    lib().orc_tridiag_eigh(m, _p(diag), _p(off), _p(w), _p(V))
    return w, V


# ---- the reference's own basis code (built from /root/reference sources) ----
_ref = {}


def ref_available(spinful=False):
    return os.path.exists(REF_SPINFUL if spinful else REF_SPINLESS)


def ref_lib(spinful=False):
    key = bool(spinful)
    if key not in _ref:
        L = C.CDLL(REF_SPINFUL if spinful else REF_SPINLESS)
        L.ref_basis.restype = C.c_int64
        L.ref_basis.argtypes = [C.c_int] * 6 + [C.c_void_p, C.c_void_p, C.c_int64]
        L.ref_prepare.argtypes = [C.c_int] * 6
        L.ref_translate.restype = C.c_uint64
        L.ref_translate.argtypes = [C.c_uint64, C.c_int, C.POINTER(C.c_int)]
        L.ref_inv_translate.restype = C.c_uint64
        L.ref_inv_translate.argtypes = [C.c_uint64, C.c_int, C.POINTER(C.c_int)]
        L.ref_sign2.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_int]
        L.ref_sign1.argtypes = [C.c_uint64, C.c_int, C.c_int]
        _ref[key] = L
    return _ref[key]


def ref_basis(sec):
    L = ref_lib(sec.S == 4)
    cap = 1 << 16
    while True:
        ids = np.empty(cap, np.uint64)
        per = np.empty(cap, np.int16)
        n = L.ref_basis(sec.nphi, sec.nel, sec.nel_up, sec.nel_su, sec.J, sec.K, _p(ids), _p(per), cap)
        if n >= 0:
            return ids[:n].copy(), per[:n].copy()
        cap = -n


def observables(basis, psi):
    """The reference's observables of psi in `basis` (dict)."""
    nphi = basis.sec.nphi
    psi = np.ascontiguousarray(psi, np.complex128)
    out = np.empty(6 + 2 * nphi)
    lib().orc_observables(basis.h, _p(psi), _p(out))
    return {"pseudospin_sz": out[0], "pseudospin_sx": out[1], "sz": out[2], "upper_sz": out[3], "lower_sz": out[4],
            "occupation": out[5:5 + 2 * nphi].reshape(2, nphi), "spinflip": out[5 + 2 * nphi]}


class planted:
    """Context: the oracle builds Hamiltonians with a planted mistake (tests
    that show a check fails on it): 1 no spinful (0,1) pair block, 2 the
    interlayer table for the (2,3) block, 3 conjugated FT phase, 4 1/sqrt(D_l)
    in place of 1/sqrt(D_l D_r); monolayer: 5 the pair sign counts empty
    orbitals instead of occupied ones, 6 conjugated FT phase."""

    def __init__(self, which):
        self.which = which

    def __enter__(self):
        lib().orc_plant(self.which)

    def __exit__(self, *a):
        lib().orc_plant(0)


def monolayer(nphi, nel, J=-1, K=-1, gamma=1.0, nLL=0, lx=None, ly=None):
    """The (J, K) sector of the single-layer Hamiltonian in the Monolayer
    tree's conventions (K < 0: no translation reduction).  The torus is
    lx x ly with lx ly = 2 pi nphi and lx / ly = gamma unless lx, ly are
    given.  Returns a dict with the basis (id, period), the CSR (rowptr, col,
    val) and E_cl."""
    if lx is None or ly is None:
        lx, ly = geometry(nphi, gamma)
    cap = 1
    while True:
        ids = np.empty(cap, np.uint64)
        per = np.empty(cap, np.int16)
        n = lib().orc_monolayer_basis(nphi, nel, J, K, _p(ids), _p(per), cap)
        if n >= 0:
            break
        cap = -n
    ecl = C.c_double()
    h = lib().orc_monolayer_new(nphi, nel, J, K, lx, ly, nLL, C.byref(ecl))
    nr = lib().orc_hamil_n(h)
    nnz = lib().orc_hamil_nnz(h)
    rp = np.empty(nr + 1, np.int64)
    col = np.empty(nnz, np.int64)
    val = np.empty(nnz, np.complex128)
    lib().orc_hamil_export(h, _p(rp), _p(col), _p(val))
    lib().orc_hamil_free(h)
    assert nr == n
    return dict(id=ids[:n].copy(), period=per[:n].copy(), rowptr=rp, col=col, val=val, ecl=ecl.value,
                lx=lx, ly=ly)