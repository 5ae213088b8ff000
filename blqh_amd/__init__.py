"""blqh_amd -- MI355X exact-diagonalisation engine for bilayer quantum Hall
systems on the torus (Python side of the C ABI in include/blqh.h).

The classes mirror the reference's class API (qiongz/BLQH):
``Basis``  ~ ``basis`` (BLQH_without_spin/basis.h:25-56),
``Hamiltonian`` ~ ``lhamil`` set_hamil / H (lanczos_hamiltonian.h:11-75),
``Hamiltonian.lanczos`` ~ coeff_explicit_update + diag +
eigenstates_reconstruction + ground_state_energy.

All compute runs in libblqh.so (HIP kernels for gfx950).  Importing this
package does not need a GPU; calling into it does, and it raises when the
library has not been built.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libblqh.so")
CLI_PATH = os.path.join(HERE, "bin", "blqh")

BLQH_OK = 0


class BlqhError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"blqh error {code}: {msg}")
        self.code = code


class Sector(C.Structure):
    _fields_ = [("nspecies", C.c_int32), ("nphi", C.c_int32), ("nel", C.c_int32), ("nel_up", C.c_int32),
                ("nel_su", C.c_int32), ("J", C.c_int32), ("K", C.c_int32)]


class Params(C.Structure):
    _fields_ = [("lx", C.c_double), ("ly", C.c_double), ("nLL", C.c_int32), ("d", C.c_double),
                ("delta_sas", C.c_double), ("delta_v", C.c_double), ("delta_z", C.c_double),
                ("theta_b", C.c_double), ("theta_x", C.c_double), ("theta_y", C.c_double)]


class LanczosOpts(C.Structure):
    _fields_ = [("lam", C.c_int32), ("seed", C.c_uint32), ("tol", C.c_double), ("check_every", C.c_int32),
                ("store_basis", C.c_int32), ("start", C.c_void_p), ("res_tol", C.c_double),
                ("skip_states", C.c_int32), ("start_mix", C.c_double)]


class LanczosResult(C.Structure):
    _fields_ = [("E0", C.c_double), ("E_rq", C.c_double), ("dE", C.c_double), ("iterations", C.c_int32),
                ("converged", C.c_int32), ("passes", C.c_int32), ("seconds", C.c_double),
                ("residual", C.c_double)]


class Observables(C.Structure):
    _fields_ = [("pseudospin_sz", C.c_double), ("pseudospin_sx", C.c_double), ("sz", C.c_double),
                ("upper_sz", C.c_double), ("lower_sz", C.c_double), ("occupation", (C.c_double * 64) * 2),
                ("spinflip", C.c_double)]


class MlpTrainOpts(C.Structure):
    _fields_ = [("steps", C.c_int32), ("batch", C.c_int32), ("lr", C.c_float), ("lr_final", C.c_float),
                ("seed", C.c_uint32)]


class MlpTrainResult(C.Structure):
    _fields_ = [("loss_first", C.c_double), ("loss_last", C.c_double), ("samples", C.c_int64),
                ("steps", C.c_int32), ("seconds", C.c_double)]


# name -> (restype, argtypes); every symbol here must exist in libblqh.so
_SIGS = {
    "blqh_last_error": (C.c_char_p, []),
    "blqh_init": (C.c_int, [C.c_int]),
    "blqh_basis_create": (C.c_int, [C.POINTER(Sector), C.POINTER(C.c_void_p)]),
    "blqh_basis_size": (C.c_int64, [C.c_void_p]),
    "blqh_basis_export": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "blqh_basis_destroy": (None, [C.c_void_p]),
    "blqh_hamil_build": (C.c_int, [C.c_void_p, C.POINTER(Params), C.POINTER(C.c_void_p)]),
    "blqh_hamil_build_terms": (C.c_int, [C.c_void_p, C.POINTER(Params), C.c_int32, C.POINTER(C.c_void_p)]),
    "blqh_hamil_combine": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double, C.POINTER(C.c_void_p)]),
    "blqh_hamil_prepare": (C.c_int, [C.c_void_p]),
    "blqh_hamil_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "blqh_hamil_export": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "blqh_hamil_destroy": (None, [C.c_void_p]),
    "blqh_hamil_get_spmv": (C.c_int, [C.c_void_p] + [C.POINTER(C.c_int32)] * 5),
    "blqh_hamil_set_spmv": (C.c_int, [C.c_void_p] + [C.c_int32] * 5),
    "blqh_spmv": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "blqh_spmv_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "blqh_lanczos": (C.c_int, [C.c_void_p, C.POINTER(LanczosOpts), C.POINTER(LanczosResult), C.c_void_p,
                               C.c_void_p, C.c_void_p]),
    "blqh_lanczos_noalloc": (C.c_int, [C.c_void_p, C.POINTER(LanczosOpts), C.POINTER(LanczosResult), C.c_void_p,
                                       C.c_void_p, C.c_void_p]),
    "blqh_batch_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]),
    "blqh_batch_lanczos": (C.c_int, [C.c_void_p, C.POINTER(LanczosOpts), C.c_void_p, C.c_void_p,
                                     C.POINTER(LanczosResult), C.c_void_p]),
    "blqh_batch_lanczos_continue": (C.c_int, [C.c_void_p, C.POINTER(LanczosOpts), C.c_void_p, C.c_double,
                                              C.POINTER(LanczosResult), C.c_void_p]),
    "blqh_batch_rebind": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_int32]),
    "blqh_batch_create_scaled": (C.c_int, [C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_int32,
                                           C.POINTER(C.c_double), C.c_int32, C.POINTER(C.c_void_p)]),
    "blqh_batch_rebind_scaled": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_int32]),
    "blqh_batch_get_lanes": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    "blqh_batch_set_lanes": (C.c_int, [C.c_void_p, C.c_int32]),
    "blqh_batch_destroy": (None, [C.c_void_p]),
    "blqh_mlp_create": (C.c_int, [C.c_int32, C.c_uint32, C.POINTER(C.c_void_p)]),
    "blqh_mlp_nparams": (C.c_int, [C.c_void_p]),
    "blqh_mlp_get_params": (C.c_int, [C.c_void_p, C.c_void_p]),
    "blqh_mlp_set_params": (C.c_int, [C.c_void_p, C.c_void_p]),
    "blqh_mlp_loss_grad": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.POINTER(C.c_double), C.c_void_p]),
    "blqh_mlp_train": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.POINTER(MlpTrainOpts), C.POINTER(MlpTrainResult)]),
    "blqh_mlp_start_vector": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_void_p]),
    "blqh_mlp_destroy": (None, [C.c_void_p]),
    "blqh_compute_observables": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Observables)]),
    "blqh_tridiag_spectrum": (C.c_int, [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "blqh_tridiag_vectors": (C.c_int, [C.c_int32, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p]),
    "blqh_lanczos_reconstruct": (C.c_int, [C.c_void_p, C.c_int32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                           C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "blqh_spectrum_distinct": (C.c_int, [C.c_int32, C.c_void_p, C.c_int32, C.c_double, C.c_void_p]),
    "blqh_record_save": (C.c_int, [C.c_char_p, C.c_int64, C.c_int32, C.c_uint32] + [C.c_void_p] * 5),
    "blqh_record_load": (C.c_int, [C.c_char_p, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                   C.POINTER(C.c_uint32)] + [C.c_void_p] * 5),
    "blqh_dense_ground_state": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_void_p]),
    "blqh_start_vector": (C.c_int, [C.c_int64, C.c_uint32, C.c_void_p]),
    "blqh_tridiag_lowest": (C.c_int, [C.c_int32, C.c_void_p, C.c_void_p, C.POINTER(C.c_double), C.c_void_p]),
    "blqh_device_copy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64]),
    "blqh_device_alloc": (C.c_int, [C.POINTER(C.c_void_p), C.c_int64]),
    "blqh_device_free": (C.c_int, [C.c_void_p]),
    "blqh_synchronize": (C.c_int, []),
    "blqh_release_cached": (C.c_int, [C.POINTER(C.c_int64)]),
    "blqh_profile_enable": (C.c_int, [C.c_int]),
    "blqh_profile_read": (C.c_int, [C.POINTER(C.c_int64), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "blqh_profile_reset": (C.c_int, []),
}


def copy_calibration(nbytes, reps=5):
    """Run the plain copy kernel `reps` times over nbytes (counter calibration)."""
    a = C.c_void_p()
    b = C.c_void_p()
    _check(lib().blqh_device_alloc(C.byref(a), nbytes))
    _check(lib().blqh_device_alloc(C.byref(b), nbytes))
    try:
        for _ in range(reps):
            _check(lib().blqh_device_copy(a, b, nbytes))
        _check(lib().blqh_synchronize())
    finally:
        lib().blqh_device_free(a)
        lib().blqh_device_free(b)


def release_cached():
    """Return the device blocks the library keeps for reuse; bytes released."""
    b = C.c_int64()
    _check(lib().blqh_release_cached(C.byref(b)))
    return b.value


def profile_enable(on=True):
    _check(lib().blqh_profile_enable(1 if on else 0))


def profile_reset():
    _check(lib().blqh_profile_reset())


def profile_read():
    """(launches, summed kernel ms, summed algorithmic bytes) of the fused Lanczos SpMV."""
    n = C.c_int64()
    ms = C.c_double()
    b = C.c_double()
    _check(lib().blqh_profile_read(C.byref(n), C.byref(ms), C.byref(b)))
    return n.value, ms.value, b.value

_lib = None


def lib():
    """Load libblqh.so; raise if it has not been built (no fallback path exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise BlqhError(-3, f"{LIB_PATH} is missing; run `python -m blqh_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != BLQH_OK:
        raise BlqhError(rc, lib().blqh_last_error().decode())


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def geometry(nphi, gamma=1.0):
    """ly = sqrt(2 pi nphi / gamma), lx = gamma ly (BLQH_without_spin/main.cpp:20-21)."""
    ly = np.sqrt(nphi * 2.0 * np.pi / gamma)
    return ly * gamma, ly


def params(nphi, gamma=1.0, nLL=0, d=0.0, delta_sas=0.0, delta_v=0.0, delta_z=0.0, theta_b=0.0, theta_x=0.0,
           theta_y=0.0):
    lx, ly = geometry(nphi, gamma)
    return Params(lx, ly, nLL, d, delta_sas, delta_v, delta_z, theta_b, theta_x, theta_y)


class Basis:
    """(J,K) sector basis: orbit representatives ``id`` (ascending) and
    periods ``period`` (the reference's id[] and basis_C[]).  nspecies 2 is
    the spinless bilayer, 4 the spinful one, 1 the monolayer (Monolayer/basis.cpp:
    translations by nphi / gcd(nel, nphi) orbitals, K < gcd(nel, nphi), no nel_up)."""

    def __init__(self, nphi, nel, nel_up=-1, J=-1, K=-1, nspecies=2, nel_su=-1):
        self.sector = Sector(nspecies, nphi, nel, nel_up, nel_su, J, K)
        self._lib = lib()
        h = C.c_void_p()
        _check(self._lib.blqh_basis_create(C.byref(self.sector), C.byref(h)))
        self.h = h
        self.n = lib().blqh_basis_size(self.h)
        self._id = None
        self._period = None

    def _export(self):
        self._id = np.empty(self.n, np.uint64)
        self._period = np.empty(self.n, np.int16)
        _check(lib().blqh_basis_export(self.h, _p(self._id), _p(self._period)))

    @property
    def id(self):
        if self._id is None:
            self._export()
        return self._id

    @property
    def period(self):
        if self._period is None:
            self._export()
        return self._period

    def __len__(self):
        return self.n

    def close(self):
        if getattr(self, "h", None):
            self._lib.blqh_basis_destroy(self.h)
            self.h = None

    def __del__(self):
        # the library handle is kept on the object: module globals may already
        # be torn down when this runs at interpreter exit
        if getattr(self, "_lib", None) is not None:
            self.close()


class Hamiltonian:
    """Hamiltonian of one sector in CSR on the GPU (lhamil::set_hamil)."""

    def __init__(self, basis, par, terms=3, _handle=None):
        """terms 3: the Hamiltonian; 1: without the interlayer tunnelling; 2: the
        tunnelling alone at Delta_SAS = 1 (for combine)."""
        self.basis = basis
        self.par = par
        self._lib = lib()
        if _handle is None:
            h = C.c_void_p()
            if terms == 3:
                _check(self._lib.blqh_hamil_build(basis.h, C.byref(par), C.byref(h)))
            else:
                _check(self._lib.blqh_hamil_build_terms(basis.h, C.byref(par), terms, C.byref(h)))
        else:
            h = _handle
        self.h = h
        n = C.c_int64()
        nnz = C.c_int64()
        _check(lib().blqh_hamil_info(self.h, C.byref(n), C.byref(nnz)))
        self.n, self.nnz = n.value, nnz.value

    @classmethod
    def combine(cls, base, tun, delta_sas):
        """H(Delta_SAS) = base + Delta_SAS * tun (base without tunnelling, tun a
        terms=2 build of the same basis and parameters)."""
        h = C.c_void_p()
        _check(lib().blqh_hamil_combine(base.h, tun.h, delta_sas, C.byref(h)))
        par = Params.from_buffer_copy(base.par)
        par.delta_sas = delta_sas
        return cls(base.basis, par, _handle=h)

    def prepare(self):
        """Build the SpMV formats now (a terms-1 or terms-2 part otherwise
        builds them on its first SpMV, solve or batch)."""
        _check(lib().blqh_hamil_prepare(self.h))
        return self

    def export(self):
        rowptr = np.empty(self.n + 1, np.int64)
        col = np.empty(self.nnz, np.int32)
        val = np.empty(self.nnz, np.complex128)
        _check(lib().blqh_hamil_export(self.h, _p(rowptr), _p(col), _p(val)))
        return rowptr, col, val

    def get_spmv(self):
        """(kind, lanes, threads, wgs_per_cu, rows) of the SpMV variant in use (kind -1: not tuned yet)."""
        v = [C.c_int32() for _ in range(5)]
        _check(lib().blqh_hamil_get_spmv(self.h, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def set_spmv(self, kind, lanes=16, threads=1024, wgs_per_cu=1, rows=2):
        _check(lib().blqh_hamil_set_spmv(self.h, kind, lanes, threads, wgs_per_cu, rows))

    def spmv(self, x):
        x = np.ascontiguousarray(x, np.complex128).reshape(-1)
        if x.size != self.n:
            raise ValueError(f"x has {x.size} entries, the Hamiltonian {self.n} rows")
        y = np.empty_like(x)
        _check(lib().blqh_spmv(self.h, _p(x), _p(y)))
        return y

    def lanczos(self, lam=200, seed=0, tol=1e-10, res_tol=1e-7, check_every=1, store_basis=-1, start=None,
                want_psi=True, skip_states=False, start_mix=0.0):
        """Ground state.  Stops when the lowest Ritz value moves by < tol and the
        Ritz residual is < res_tol (tol <= 0: exactly lam steps).  skip_states:
        passes 1 and 2 only (alpha, beta, E0; no psi, E_rq is NaN).  start_mix:
        weight of the seeded random unit vector added to a given start."""
        return self._lanczos(lib().blqh_lanczos, lam, seed, tol, res_tol, check_every, store_basis, start,
                             want_psi, skip_states, start_mix)

    def lanczos_noalloc(self, lam=200, seed=0, tol=1e-10, res_tol=1e-7, check_every=1, start=None, want_psi=True,
                        skip_states=False):
        """lanczos() through blqh_lanczos_noalloc: no device allocation (the
        solver's static per-GPU scratch; at most 131072 rows, lam <= 1024),
        three passes as lanczos(store_basis=0)."""
        return self._lanczos(lib().blqh_lanczos_noalloc, lam, seed, tol, res_tol, check_every, 0, start, want_psi,
                             skip_states, 0.0)

    def _lanczos(self, fn, lam, seed, tol, res_tol, check_every, store_basis, start, want_psi, skip_states,
                 start_mix):
        opts = LanczosOpts(lam, seed, tol, check_every, store_basis, None, res_tol, 1 if skip_states else 0,
                           start_mix)
        keep = None
        if start is not None:
            keep = np.ascontiguousarray(start, np.complex128).reshape(-1)
            if keep.size != self.n:
                raise ValueError(f"start vector has {keep.size} entries, the Hamiltonian {self.n} rows")
            opts.start = keep.ctypes.data
        res = LanczosResult()
        psi = np.empty(self.n, np.complex128) if want_psi else None
        a = np.zeros(lam)
        b = np.zeros(lam + 1)
        _check(fn(self.h, C.byref(opts), C.byref(res), _p(psi) if want_psi else None, _p(a), _p(b)))
        out = {k: getattr(res, k) for k, _ in LanczosResult._fields_}
        out.update(psi=psi, alpha=a[:res.iterations], beta=b[:res.iterations + 1])
        return out

    def reconstruct(self, coefs, alpha=None, beta=None, seed=0, start=None):
        """Pass 3 (eigenstates_reconstruction): regenerate the Lanczos vectors of
        len(coefs[0]) steps from `seed` (or `start`) and return the normalised
        states psi_s = sum_l coefs[s][l] conj(v_l), their Rayleigh quotients and
        the largest relative deviation of the regenerated alpha / beta from the
        given ones (beta[0] = 1 as returned by lanczos())."""
        c = np.ascontiguousarray(np.atleast_2d(coefs), np.float64)
        ns, m = c.shape
        a = None if alpha is None else np.ascontiguousarray(alpha, np.float64).reshape(-1)
        b = None if beta is None else np.ascontiguousarray(beta, np.float64).reshape(-1)
        if a is not None and len(a) < m - 1:
            raise ValueError(f"alpha has {len(a)} entries, {m - 1} needed")
        if b is not None and len(b) < m:
            raise ValueError(f"beta has {len(b)} entries, {m} needed")
        if not 1 <= ns <= 16:
            raise ValueError(f"{ns} coefficient vectors (1 to 16 allowed)")
        keep = None if start is None else np.ascontiguousarray(start, np.complex128).reshape(-1)
        if keep is not None and keep.size != self.n:
            raise ValueError(f"start vector has {keep.size} entries, the Hamiltonian {self.n} rows")
        psi = np.empty((ns, self.n), np.complex128)
        erq = np.empty(ns)
        dev = C.c_double()
        _check(lib().blqh_lanczos_reconstruct(self.h, m, seed, _p(keep) if keep is not None else None,
                                              _p(a) if a is not None else None, _p(b) if b is not None else None,
                                              ns, _p(c), _p(psi), _p(erq), C.byref(dev)))
        return psi, erq, dev.value

    def eigenstates(self, lam=200, seed=0, tol=1e-10, res_tol=1e-7, check_every=1, start=None, e1_gap=0.0):
        """The spinful reference's three passes with the first excited state
        (coeff_explicit_update, diag, eigenstates_reconstruction,
        ground_state_energy, first_excited_state_energy): psi_0 and psi_1 are
        rebuilt in one replay.  The excited Krylov vector is eigenvector 1 of
        the tridiagonal, as the reference's diag() selects in effect (its
        1e-2 gap test reads the eigenvalues before they are filled); e1_gap > 0
        picks the first eigenvalue more than e1_gap above E0 instead."""
        r = self.lanczos(lam, seed, tol, res_tol, check_every, 0, start, want_psi=False, skip_states=True)
        a, b = r["alpha"], r["beta"]
        m = len(a)
        if m < 2:
            raise BlqhError(-1, "the first excited state needs at least 2 Lanczos vectors")
        eig, _, _ = tridiag_spectrum(a, b)
        k1 = 1
        if e1_gap > 0:
            big = np.nonzero(np.abs(eig - eig[0]) > e1_gap)[0]
            if len(big) == 0:
                raise BlqhError(-1, "no eigenvalue above E0 + e1_gap in the Krylov space")
            k1 = int(big[0])
        v = tridiag_vectors(a, b, 0, k1 + 1)[[0, k1]]
        psi, erq, dev = self.reconstruct(v, a, b, seed=seed, start=start)
        r.update(psi=psi[0], psi1=psi[1], E_rq=erq[0], E1_rq=erq[1], E1=eig[k1], e1_index=k1, replay_dev=dev,
                 eigenvalues=eig)
        return r

    def dense_ground_state(self):
        e0 = C.c_double()
        ep = C.c_double()
        v = np.empty(self.n, np.complex128)
        _check(lib().blqh_dense_ground_state(self.h, C.byref(e0), C.byref(ep), _p(v)))
        return e0.value, ep.value, v

    def close(self):
        if getattr(self, "h", None):
            self._lib.blqh_hamil_destroy(self.h)
            self.h = None

    def __del__(self):
        # the library handle is kept on the object: module globals may already
        # be torn down when this runs at interpreter exit
        if getattr(self, "_lib", None) is not None:
            self.close()


class Batch:
    """Resident Hamiltonians solved together (blqh_batch_*).  The Hamiltonian
    objects are kept alive by the batch."""

    def __init__(self, hams, lanes=0):
        """lanes=0 autotunes the batched SpMV; otherwise 4, 8, 16, 32 or 64."""
        self.hams = list(hams)
        self._lib = lib()
        arr = (C.c_void_p * len(self.hams))(*[H.h.value for H in self.hams])
        h = C.c_void_p()
        _check(self._lib.blqh_batch_create(arr, len(self.hams), lanes, C.byref(h)))
        self.h = h

    @classmethod
    def scaled(cls, bases, tuns, scales):
        """Shared-scale batch (blqh_batch_create_scaled): problem q * len(scales) + g
        is bases[q] + scales[g] tuns[q] (bases built with terms=1, tuns with terms=2).
        The batch keeps its own copy of the matrices; ``hams`` lists, per
        problem, the sector's base (its size and basis)."""
        bases, tuns, scales = list(bases), list(tuns), [float(x) for x in scales]
        if len(bases) != len(tuns) or not bases:
            raise ValueError(f"{len(bases)} base and {len(tuns)} tunnelling matrices")
        if not 1 <= len(scales) <= 4:
            raise ValueError(f"{len(scales)} scales (1 to 4)")
        self = cls.__new__(cls)
        self._lib = lib()
        a = (C.c_void_p * len(bases))(*[H.h.value for H in bases])
        t = (C.c_void_p * len(tuns))(*[H.h.value for H in tuns])
        sc = (C.c_double * len(scales))(*scales)
        h = C.c_void_p()
        _check(self._lib.blqh_batch_create_scaled(a, t, len(bases), sc, len(scales), C.byref(h)))
        self.h = h
        self.scales = scales
        self.hams = [H for H in bases for _ in scales]
        return self

    def rebind_scaled(self, bases, tuns):
        """The next d of a shared-scale batch (same sectors, same order)."""
        bases, tuns = list(bases), list(tuns)
        if len(bases) != len(tuns):
            raise ValueError(f"{len(bases)} base and {len(tuns)} tunnelling matrices")
        a = (C.c_void_p * len(bases))(*[H.h.value for H in bases])
        t = (C.c_void_p * len(tuns))(*[H.h.value for H in tuns])
        _check(self._lib.blqh_batch_rebind_scaled(self.h, a, t, len(bases)))
        self.hams = [H for H in bases for _ in self.scales]

    def rebind(self, hams):
        """Members of the same sizes, same order (next point of a sweep)."""
        hams = list(hams)
        arr = (C.c_void_p * len(hams))(*[H.h.value for H in hams])
        _check(self._lib.blqh_batch_rebind(self.h, arr, len(hams)))
        self.hams = hams

    @property
    def lanes(self):
        v = C.c_int32()
        _check(self._lib.blqh_batch_get_lanes(self.h, C.byref(v)))
        return v.value

    @lanes.setter
    def lanes(self, v):
        _check(self._lib.blqh_batch_set_lanes(self.h, v))

    def lanczos(self, lam=200, seeds=0, tol=1e-10, res_tol=1e-7, check_every=5, starts=None, want_psi=True,
                start_mix=0.0, continue_prev=False, excited_mix=0.0):
        """Ground states of every problem; `seeds` is one seed or one per problem,
        `starts` None or a list of start vectors (None entries: seeded),
        `start_mix` the weight of the seeded unit vector added to each start.
        continue_prev: start every problem from conj(psi) of the previous
        solve on this batch, kept on the device (blqh_batch_lanczos_continue;
        the same lam as that solve, no `starts`); excited_mix adds that
        weight of the previous solve's second and third Ritz vectors."""
        P = len(self.hams)
        opts = LanczosOpts(lam, 0, tol, check_every, 1, None, res_tol, 0, start_mix)
        sd = np.asarray(seeds, np.uint32)
        if sd.ndim > 1 or (sd.ndim == 1 and sd.size != P):
            raise ValueError(f"{sd.size} seeds for {P} problems")
        sd = np.broadcast_to(sd, (P,)).copy()
        keep = []
        st = None
        if starts is not None:
            starts = list(starts)
            if len(starts) != P:
                raise ValueError(f"{len(starts)} start vectors for {P} problems")
            st = (C.c_void_p * P)()
            for i, v in enumerate(starts):
                if v is not None:
                    a = np.ascontiguousarray(v, np.complex128).reshape(-1)
                    if a.size != self.hams[i].n:
                        raise ValueError(f"start vector {i} has {a.size} entries, problem {i} {self.hams[i].n} rows")
                    keep.append(a)
                    st[i] = a.ctypes.data
        res = (LanczosResult * P)()
        psis = [np.empty(H.n, np.complex128) for H in self.hams] if want_psi else None
        po = None
        if want_psi:
            po = (C.c_void_p * P)(*[x.ctypes.data for x in psis])
        if continue_prev:
            if starts is not None:
                raise ValueError("continue_prev takes no start vectors")
            _check(self._lib.blqh_batch_lanczos_continue(self.h, C.byref(opts), sd.ctypes.data, float(excited_mix),
                                                         res, po))
        else:
            _check(self._lib.blqh_batch_lanczos(self.h, C.byref(opts), sd.ctypes.data, st, res, po))
        out = []
        for i in range(P):
            d = {k: getattr(res[i], k) for k, _ in LanczosResult._fields_}
            d["psi"] = psis[i] if want_psi else None
            out.append(d)
        return out

    def close(self):
        if getattr(self, "h", None):
            self._lib.blqh_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        if getattr(self, "_lib", None) is not None:
            self.close()


class Mlp:
    """Learned warm start (blqh_mlp_*): FP32 MLP over occupation bits and
    (d, Delta_SAS) predicting log|psi| and the sign of psi."""

    def __init__(self, nbits, seed=0):
        self._lib = lib()
        h = C.c_void_p()
        _check(self._lib.blqh_mlp_create(nbits, seed, C.byref(h)))
        self.h = h
        self.nbits = nbits
        self.nparams = self._lib.blqh_mlp_nparams(self.h)

    @property
    def params(self):
        out = np.empty(self.nparams, np.float32)
        _check(self._lib.blqh_mlp_get_params(self.h, _p(out)))
        return out

    @params.setter
    def params(self, v):
        v = np.ascontiguousarray(v, np.float32).reshape(-1)
        if v.size != self.nparams:
            raise ValueError(f"{v.size} parameters, the network has {self.nparams}")
        _check(self._lib.blqh_mlp_set_params(self.h, _p(v)))

    def loss_grad(self, words, d, dsas, ya, ys):
        """Mean loss and gradient over explicit states (all arrays of one length)."""
        w = np.ascontiguousarray(words, np.uint64).reshape(-1)
        f = [np.ascontiguousarray(a, np.float32).reshape(-1) for a in (d, dsas, ya, ys)]
        if any(a.size != w.size for a in f):
            raise ValueError("words, d, dsas, ya and ys must have one length")
        loss = C.c_double()
        g = np.empty(self.nparams, np.float32)
        _check(self._lib.blqh_mlp_loss_grad(self.h, len(w), _p(w), _p(f[0]), _p(f[1]), _p(f[2]), _p(f[3]),
                                            C.byref(loss), _p(g)))
        return loss.value, g

    def train(self, sets, steps=2000, batch=4096, lr=3e-3, lr_final=3e-4, seed=1):
        """sets: list of (Basis, d, dsas, psi).  Returns a dict of the run."""
        n = len(sets)
        keep = [np.ascontiguousarray(p, np.complex128).reshape(-1) for _, _, _, p in sets]
        for k, (b, _, _, _) in enumerate(sets):
            if keep[k].size != b.n:
                raise ValueError(f"state {k} has {keep[k].size} entries, its basis {b.n} states")
        bases = (C.c_void_p * n)(*[b.h.value for b, _, _, _ in sets])
        par = np.ascontiguousarray([[d, s] for _, d, s, _ in sets], np.float64)
        psi = (C.c_void_p * n)(*[k.ctypes.data for k in keep])
        o = MlpTrainOpts(steps, batch, lr, lr_final, seed)
        r = MlpTrainResult()
        _check(self._lib.blqh_mlp_train(self.h, n, bases, _p(par), psi, C.byref(o), C.byref(r)))
        return {k: getattr(r, k) for k, _ in MlpTrainResult._fields_}

    def start_vector(self, basis, d, dsas):
        out = np.empty(basis.n, np.complex128)
        _check(self._lib.blqh_mlp_start_vector(self.h, basis.h, d, dsas, _p(out)))
        return out

    def close(self):
        if getattr(self, "h", None):
            self._lib.blqh_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        if getattr(self, "_lib", None) is not None:
            self.close()


def observables(basis, psi):
    """Ground-state observables of psi in `basis` (dict; occupation is [2][nphi])."""
    v = np.ascontiguousarray(psi, np.complex128).reshape(-1)
    if v.size != basis.n:
        raise ValueError(f"state has {v.size} entries, the basis {basis.n} states")
    o = Observables()
    _check(lib().blqh_compute_observables(basis.h, _p(v), C.byref(o)))
    nphi = basis.sector.nphi
    occ = np.array([[o.occupation[a][j] for j in range(nphi)] for a in range(2)])
    return {"pseudospin_sz": o.pseudospin_sz, "pseudospin_sx": o.pseudospin_sx, "sz": o.sz,
            "upper_sz": o.upper_sz, "lower_sz": o.lower_sz, "occupation": occ, "spinflip": o.spinflip}


def tridiag_spectrum(alpha, beta):
    """Eigenpairs of the Lanczos tridiagonal: alpha[m], beta[m+1] as returned by
    lanczos() (beta[0] = 1).  Returns (eigenvalues, psi_n0, psi_0)."""
    a = np.ascontiguousarray(alpha, np.float64)
    m = len(a)
    b = np.ascontiguousarray(beta, np.float64)[1:m]
    eig, first, vec0 = np.empty(m), np.empty(m), np.empty(m)
    _check(lib().blqh_tridiag_spectrum(m, _p(a), _p(b) if m > 1 else None, _p(eig), _p(first), _p(vec0)))
    return eig, first, vec0


def tridiag_vectors(alpha, beta, k0, count):
    """Eigenvectors k0..k0+count-1 of the Lanczos tridiagonal (alpha[m], beta[m+1],
    beta[0] = 1), rows of the returned (count, m) array, first components >= 0."""
    a = np.ascontiguousarray(alpha, np.float64)
    m = len(a)
    b = np.ascontiguousarray(beta, np.float64)[1:m]
    out = np.empty((count, m))
    _check(lib().blqh_tridiag_vectors(m, _p(a), _p(b) if m > 1 else None, k0, count, _p(out)))
    return out


def spectrum_distinct(eig, kmax=7, gap=1e-5):
    """Ghost-filtered low-lying levels (eig[0], then jumps > gap), at most kmax."""
    e = np.ascontiguousarray(eig, np.float64)
    out = np.empty(max(kmax, 1))
    n = lib().blqh_spectrum_distinct(len(e), _p(e), kmax, gap, _p(out))
    _check(0 if n >= 0 else n)
    return out[:n]


def record_save(path, nhilbert, seed, eig, beta, alpha, psi0, psin0):
    m = len(alpha)
    arrs = [np.ascontiguousarray(v, np.float64) for v in (eig, beta, alpha, psi0, psin0)]
    if len(arrs[1]) != m + 1 or any(len(v) != m for i, v in enumerate(arrs) if i != 1):
        raise ValueError("record arrays: beta needs len(alpha) + 1 entries, the others len(alpha)")
    _check(lib().blqh_record_save(os.fsencode(path), nhilbert, m, seed, *[_p(v) for v in arrs]))


def record_load(path, cap=100000):
    n, m, sd = C.c_int64(), C.c_int32(), C.c_uint32()
    eig, beta, alpha, psi0, psin0 = (np.empty(cap), np.empty(cap + 1), np.empty(cap), np.empty(cap), np.empty(cap))
    _check(lib().blqh_record_load(os.fsencode(path), cap, C.byref(n), C.byref(m), C.byref(sd), _p(eig), _p(beta),
                                  _p(alpha), _p(psi0), _p(psin0)))
    k = m.value
    return {"nhilbert": n.value, "lambda": k, "seed": sd.value, "eigenvalues": eig[:k], "norm": beta[:k + 1],
            "overlap": alpha[:k], "psi_0": psi0[:k], "psi_n0": psin0[:k]}


def start_vector(n, seed):
    out = np.empty(n, np.complex128)
    _check(lib().blqh_start_vector(n, seed, _p(out)))
    return out


def tridiag_lowest(alpha, offdiag):
    alpha = np.ascontiguousarray(alpha, np.float64)
    offdiag = np.ascontiguousarray(offdiag, np.float64)
    e0 = C.c_double()
    v = np.empty(len(alpha))
    _check(lib().blqh_tridiag_lowest(len(alpha), _p(alpha), _p(offdiag), C.byref(e0), _p(v)))
    return e0.value, v
