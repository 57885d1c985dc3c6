"""ORACLE — TEST INFRASTRUCTURE ONLY.

ctypes wrapper over oracle/_build/liboracle.so, the sequential CPU restatement
of MLD2P4 2.1's smoothed-aggregation AMG (see oracle.cpp for the per-routine
reference citations and pinning status).  Only tests/, the smoke check and the
CPU-baseline leg of the benchmark may import this module; the product library
never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

# Smoother codes: mld_base_prec_type.F90:166-176 (0..2); 100+ are additions.
SM_NOPREC, SM_JAC, SM_BJAC, SM_L1JAC, SM_DENSELU = 0, 1, 2, 100, 101
SM_FBGS, SM_GS, SM_BWGS = 4, 102, 103
SM_AS = 3
CYC_ADD, CYC_MULT, CYC_V, CYC_W, CYC_K, CYC_KSYM = 1, 2, 3, 4, 5, 6
KRYLOV = {"CG": 0, "FCG": 1, "BICGSTAB": 2}


class OrcParams(C.Structure):
    _fields_ = [
        ("theta", C.c_double),
        ("min_coarse_size", C.c_int),
        ("max_levs", C.c_int),
        ("min_cr_ratio", C.c_double),
        ("aggr_omega_alg", C.c_int),
        ("aggr_omega_val", C.c_double),
        ("aggr_filter", C.c_int),
        ("aggr_ord", C.c_int),
        ("ml_cycle", C.c_int),
        ("outer_sweeps", C.c_int),
        ("smoother", C.c_int),
        ("sweeps_pre", C.c_int),
        ("sweeps_post", C.c_int),
        ("c_smoother", C.c_int),
        ("c_sweeps", C.c_int),
        ("sub_fill", C.c_int),
        ("sub_milu", C.c_int),
        ("c_fill", C.c_int),
        ("c_milu", C.c_int),
    ]


class OrcCsr(C.Structure):
    _fields_ = [
        ("nr", C.c_int64),
        ("nc", C.c_int64),
        ("nnz", C.c_int64),
        ("irp", C.POINTER(C.c_int64)),
        ("ja", C.POINTER(C.c_int)),
        ("val", C.POINTER(C.c_double)),
    ]


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp, dp, ip, i64p = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_int64)
        L.orc_set_threads.restype = C.c_int
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_gen_pde3d.restype = vp
        L.orc_gen_pde3d.argtypes = [C.c_int]
        L.orc_csr_new.restype = vp
        L.orc_csr_new.argtypes = [C.c_int64, C.c_int64, i64p, ip, dp]
        L.orc_csr_free.argtypes = [vp]
        L.orc_csr_view.argtypes = [vp, C.POINTER(OrcCsr)]
        L.orc_spmv.argtypes = [vp, C.c_double, dp, C.c_double, dp]
        L.orc_spmv_t.argtypes = [vp, C.c_double, dp, C.c_double, dp]
        L.orc_dec_map_bld.argtypes = [vp, C.c_int, C.c_double, ip, ip]
        L.orc_ilu0.argtypes = [vp, C.POINTER(vp), C.POINTER(vp), dp]
        L.orc_trsv.argtypes = [vp, C.c_int, C.c_int, dp, dp]
        L.orc_iluk.argtypes = [vp, C.c_int, C.c_int, C.POINTER(vp), C.POINTER(vp), dp]
        L.orc_prec_new.restype = vp
        L.orc_prec_new.argtypes = [C.POINTER(OrcParams)]
        L.orc_prec_free.argtypes = [vp]
        L.orc_prec_build.argtypes = [vp, vp]
        L.orc_prec_nlev.argtypes = [vp]
        L.orc_prec_casize.argtypes = [vp]
        L.orc_prec_omega.restype = C.c_double
        L.orc_prec_omega.argtypes = [vp, C.c_int]
        L.orc_prec_level_csr.restype = vp
        L.orc_prec_level_csr.argtypes = [vp, C.c_int, C.c_int]
        L.orc_prec_level_naggr.argtypes = [vp, C.c_int]
        L.orc_prec_level_ilaggr.argtypes = [vp, C.c_int, ip]
        L.orc_prec_level_dv.argtypes = [vp, C.c_int, dp]
        L.orc_prec_apply.argtypes = [vp, dp, dp, C.c_char]
        L.orc_prec_smooth.argtypes = [vp, C.c_int, C.c_double, dp, C.c_double, dp, C.c_char, C.c_int]
        L.orc_prec_smooth_w.argtypes = [vp, C.c_int, C.c_int, C.c_double, dp, C.c_double, dp, C.c_char, C.c_int]
        L.orc_prec_level_cycle.argtypes = [vp, C.c_int, dp, dp, C.c_char]
        L.orc_krylov.argtypes = [vp, vp, C.c_int, dp, dp, C.c_double, C.c_int, ip, dp]
        L.orc_krylov_timed.argtypes = [vp, vp, C.c_int, dp, dp, C.c_double, C.c_int, ip, dp, dp, dp]
        L.orc_prec_build_timed.argtypes = [vp, vp, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class Csr:
    """A CSR matrix owned by the oracle (0-based, int64 row pointers)."""

    def __init__(self, handle, owned=True):
        self.h = C.c_void_p(handle)
        self.owned = owned

    @classmethod
    def from_arrays(cls, irp, ja, val, ncols=None):
        irp = np.ascontiguousarray(irp, dtype=np.int64)
        ja = np.ascontiguousarray(ja, dtype=np.int32)
        val = np.ascontiguousarray(val, dtype=np.float64)
        n = len(irp) - 1
        nc = n if ncols is None else ncols
        h = lib().orc_csr_new(n, nc, irp.ctypes.data_as(C.POINTER(C.c_int64)), _ip(ja), _dp(val))
        return cls(h)

    @classmethod
    def from_scipy(cls, m):
        m = m.tocsr()
        m.sort_indices()
        return cls.from_arrays(m.indptr, m.indices, m.data, ncols=m.shape[1])

    def __del__(self):
        if self.owned and self.h and _lib is not None:
            _lib.orc_csr_free(self.h)
            self.h = None

    def arrays(self):
        v = OrcCsr()
        lib().orc_csr_view(self.h, C.byref(v))
        nr, nnz = v.nr, v.nnz
        irp = np.ctypeslib.as_array(v.irp, shape=(nr + 1,)).copy() if nr >= 0 else np.zeros(1, np.int64)
        ja = np.ctypeslib.as_array(v.ja, shape=(nnz,)).copy() if nnz else np.zeros(0, np.int32)
        val = np.ctypeslib.as_array(v.val, shape=(nnz,)).copy() if nnz else np.zeros(0)
        return irp, ja, val, (v.nr, v.nc)

    def to_scipy(self):
        import scipy.sparse as sp

        irp, ja, val, shape = self.arrays()
        return sp.csr_matrix((val, ja, irp), shape=shape)

    @property
    def shape(self):
        v = OrcCsr()
        lib().orc_csr_view(self.h, C.byref(v))
        return (v.nr, v.nc)

    def spmv(self, x, alpha=1.0, beta=0.0, y=None, trans=False):
        x = np.ascontiguousarray(x, dtype=np.float64)
        nr, nc = self.shape
        if y is None:
            y = np.zeros(nc if trans else nr)
        (lib().orc_spmv_t if trans else lib().orc_spmv)(self.h, alpha, _dp(x), beta, _dp(y))
        return y


def set_threads(n=0):
    """Threads of the oracle's row loops (0: OpenMP's default); returns the
    count in use.  Results do not depend on it."""
    return lib().orc_set_threads(int(n))


def gen_pde3d(idim):
    """7-point 3-D Poisson of tests/pdegen/mld_d_pde3d.f90:98-412."""
    return Csr(lib().orc_gen_pde3d(idim))


def dec_map_bld(a: Csr, theta=0.0, iorder=0):
    n = a.shape[0]
    ilaggr = np.zeros(n, dtype=np.int32)
    naggr = C.c_int(0)
    rc = lib().orc_dec_map_bld(a.h, iorder, theta, _ip(ilaggr), C.byref(naggr))
    if rc:
        raise RuntimeError(f"dec_map_bld failed: {rc}")
    return ilaggr, naggr.value


def ilu0(a: Csr):
    n = a.shape[0]
    d = np.zeros(n)
    lh, uh = C.c_void_p(), C.c_void_p()
    rc = lib().orc_ilu0(a.h, C.byref(lh), C.byref(uh), _dp(d))
    return Csr(lh.value), Csr(uh.value), d, rc


def iluk(a: Csr, fill, milu=False):
    n = a.shape[0]
    d = np.zeros(n)
    lh, uh = C.c_void_p(), C.c_void_p()
    rc = lib().orc_iluk(a.h, int(fill), int(milu), C.byref(lh), C.byref(uh), _dp(d))
    return Csr(lh.value), Csr(uh.value), d, rc


def trsv(t: Csr, x, lower=True, trans=False):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros_like(x)
    lib().orc_trsv(t.h, int(lower), int(trans), _dp(x), _dp(y))
    return y


@dataclass
class Params:
    theta: float = 0.01            # mld_d_onelev_mod.f90:446 default
    min_coarse_size: int = -1      # library default 40*floor(N^(1/3))
    max_levs: int = 20             # mld_d_prec_type.f90:92
    min_cr_ratio: float = 1.5
    aggr_omega_alg: int = 0
    aggr_omega_val: float = 0.0
    aggr_filter: int = 0
    aggr_ord: int = 0
    ml_cycle: int = CYC_V
    outer_sweeps: int = 1
    smoother: int = SM_JAC
    sweeps_pre: int = 1
    sweeps_post: int = 1
    c_smoother: int = SM_BJAC
    c_sweeps: int = 1
    sub_fill: int = 0
    sub_milu: int = 0
    c_fill: int = 0
    c_milu: int = 0

    def to_c(self):
        return OrcParams(**{f: getattr(self, f) for f, _ in OrcParams._fields_})


class Prec:
    """Oracle preconditioner: hierarchy build + V-cycle apply + Krylov."""

    def __init__(self, params: Params):
        self.params = params
        self._p = params.to_c()
        self.h = C.c_void_p(lib().orc_prec_new(C.byref(self._p)))
        self.a = None

    def __del__(self):
        if self.h and _lib is not None:
            _lib.orc_prec_free(self.h)
            self.h = None

    def build(self, a: Csr):
        self.a = a
        t = C.c_double(0.0)
        rc = lib().orc_prec_build_timed(self.h, a.h, C.byref(t))
        if rc:
            raise RuntimeError(f"oracle build failed: {rc}")
        self.build_s = t.value  # hierarchy + smoothers, steady_clock seconds
        return self

    @property
    def nlev(self):
        return lib().orc_prec_nlev(self.h)

    @property
    def casize(self):
        return lib().orc_prec_casize(self.h)

    def omega(self, lev):
        return lib().orc_prec_omega(self.h, lev)

    def level_csr(self, lev, what):
        """what: 'A', 'P', 'R', 'ND', 'L', 'U' (borrowed view, copied out)."""
        code = {"A": 0, "P": 1, "R": 2, "ND": 3, "L": 4, "U": 5}[what]
        return Csr(lib().orc_prec_level_csr(self.h, lev, code), owned=False)

    def naggr(self, lev):
        return lib().orc_prec_level_naggr(self.h, lev)

    def ilaggr(self, lev):
        n = self.level_csr(lev - 1, "A").shape[0]
        out = np.zeros(n, dtype=np.int32)
        lib().orc_prec_level_ilaggr(self.h, lev, _ip(out))
        return out

    def dv(self, lev):
        n = self.level_csr(lev, "A").shape[0]
        out = np.zeros(n)
        lib().orc_prec_level_dv(self.h, lev, _dp(out))
        return out

    def apply(self, x, trans="N"):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        rc = lib().orc_prec_apply(self.h, _dp(x), _dp(y), trans.encode())
        if rc:
            raise RuntimeError(f"oracle apply failed: {rc}")
        return y

    def level_cycle(self, lev, x, trans="N"):
        """The cycle from level lev down applied to x (inner_ml_aply)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        rc = lib().orc_prec_level_cycle(self.h, lev, _dp(x), _dp(y), trans.encode())
        if rc:
            raise RuntimeError(f"oracle level cycle failed: {rc}")
        return y

    def smooth(self, lev, x, alpha=1.0, beta=0.0, y=None, trans="N", sweeps=1, which=0):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x) if y is None else np.ascontiguousarray(y, dtype=np.float64).copy()
        rc = lib().orc_prec_smooth_w(self.h, lev, which, alpha, _dp(x), beta, _dp(y), trans.encode(), sweeps)
        if rc:
            raise RuntimeError(f"oracle smooth failed: {rc}")
        return y

    def krylov(self, method, b, x0=None, eps=1e-6, itmax=500, timings=False):
        """(x, iterations, relres); with timings=True also (per-iteration
        seconds, whole-call seconds), both from steady_clock."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b) if x0 is None else np.ascontiguousarray(x0, dtype=np.float64).copy()
        it = C.c_int(0)
        err = C.c_double(0.0)
        its = np.zeros(max(itmax, 1))
        tot = C.c_double(0.0)
        rc = lib().orc_krylov_timed(self.h, self.a.h, KRYLOV[method.upper()], _dp(b), _dp(x), eps, itmax,
                                    C.byref(it), C.byref(err), _dp(its), C.byref(tot))
        if rc:
            raise RuntimeError(f"oracle krylov failed: {rc}")
        if timings:
            return x, it.value, err.value, its[:it.value].copy(), tot.value
        return x, it.value, err.value


def config1_params(**kw):
    """Configuration of the idim=40 CPU case: JACOBI 1/1, theta 0, smoothed,
    MIN_COARSE_SIZE 200, BJAC/ILU(0) coarse with 4 sweeps."""
    p = Params(theta=0.0, min_coarse_size=200, smoother=SM_JAC, sweeps_pre=1, sweeps_post=1,
               c_smoother=SM_BJAC, c_sweeps=4)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def read_mtx(path):
    """MatrixMarket reader (coordinate real general / array real general)."""
    import scipy.sparse as sp

    with open(path) as f:
        header = f.readline().lower().split()
        line = f.readline()
        while line.startswith("%"):
            line = f.readline()
        dims = [int(t) for t in line.split()]
        data = np.loadtxt(f, ndmin=1)
    if header[2] == "coordinate":
        nr, nc, nz = dims
        data = data.reshape(nz, -1)
        rows = data[:, 0].astype(np.int64) - 1
        cols = data[:, 1].astype(np.int64) - 1
        m = sp.coo_matrix((data[:, 2], (rows, cols)), shape=(nr, nc)).tocsr()
        m.sum_duplicates()
        m.sort_indices()
        return m
    nr, nc = dims
    return data.reshape(nc, nr).T.copy() if nc > 1 else data.reshape(nr)
