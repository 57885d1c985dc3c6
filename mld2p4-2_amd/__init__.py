"""Python host mirror of MLD2P4's preconditioner interface over the MI355X engine.

The class ``MldPrec`` keeps the reference's method names and argument meaning
(``mlprec/mld_d_prec_type.f90:84-142``):

    p = MldPrec()                 # type(mld_dprec_type) :: p
    p.init('ML')                  # call p%init('ML', info)
    p.set('SMOOTHER_TYPE', 'JACOBI')
    p.hierarchy_build(a)          # call p%hierarchy_build(a, desc, info)
    p.smoothers_build()           # call p%smoothers_build(a, desc, info)
    p.apply(x, y)                 # call p%apply(x, y, desc, info)
    x, it, err = psb_krylov('CG', a, p, b, x0, eps=1e-6)

Errors are raised as ``MldError`` carrying the reference's ``info`` code.
The compute path is the HIP library ``_build/libmld2p4_amd.so``; there is no
CPU fallback: without the library (or without a GPU for apply/krylov) the
calls fail loudly.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libmld2p4_amd.so")
_lib = None


class MldError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"info={code}: {msg}")
        self.info = code


def build(jobs=8):
    """Compile the HIP library for gfx950 (make -C this directory)."""
    subprocess.run(["make", "-s", "-j", str(jobs), "-C", _HERE], check=True)


def lib():
    """Load the engine library (it must have been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MldError(-30, f"engine library not built: {LIB_PATH} (run build())")
        L = C.CDLL(LIB_PATH)
        vp, cp, dp, ip, i64p = C.c_void_p, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_int64)
        sig = {
            "mldhip_create": [C.POINTER(vp), C.c_int],
            "mldhip_destroy": [vp],
            "mldhip_init": [vp, cp],
            "mldhip_seti": [vp, cp, C.c_int, C.c_int, C.c_int, cp],
            "mldhip_setc": [vp, cp, cp, C.c_int, C.c_int, cp],
            "mldhip_setr": [vp, cp, C.c_double, C.c_int, C.c_int, cp],
            "mldhip_set_precision": [vp, C.c_int],
            "mldhip_set_format": [vp, C.c_int],
            "mldhip_set_setup": [vp, C.c_int],
            "mldhip_hierarchy_build": [vp, C.c_int, i64p, ip, dp],
            "mldhip_smoothers_build": [vp],
            "mldhip_build": [vp, C.c_int, i64p, ip, dp],
            "mldhip_upload_level": [vp, C.c_int, C.c_int, i64p, ip, dp, C.c_int, i64p, ip, dp, i64p, ip, dp],
            "mldhip_upload_smoother": [vp, C.c_int, C.c_int, C.c_int, i64p, ip, dp, dp, i64p, ip, dp, i64p, ip, dp],
            "mldhip_finalize_hierarchy": [vp, C.c_int],
            "mldhip_load_hierarchy": [vp, cp, C.c_int, i64p, ip, dp, C.c_int],
            "mldhip_mlprec_aply": [vp, C.c_double, vp, C.c_double, vp, C.c_char],
            "mldhip_apply": [vp, vp, vp, C.c_char],
            "mldhip_apply_host": [vp, dp, dp, C.c_char],
            "mldhip_krylov": [vp, cp, vp, vp, C.c_double, C.c_int, C.c_int, ip, dp],
            "mldhip_krylov_host": [vp, cp, dp, dp, C.c_double, C.c_int, C.c_int, ip, dp],
            "mldhip_krylov_noalloc": [vp, cp, vp, vp, C.c_double, C.c_int, C.c_int, ip, dp],
            "mldhip_set_stream": [vp, vp],
            "mldhip_synchronize": [vp],
            "mldhip_get_nlev": [vp, ip],
            "mldhip_get_level_info": [vp, C.c_int, i64p],
            "mldhip_get_csr_size": [vp, C.c_int, C.c_int, i64p, i64p, i64p],
            "mldhip_get_level_csr": [vp, C.c_int, C.c_int, i64p, ip, dp],
            "mldhip_get_aggr": [vp, C.c_int, ip],
            "mldhip_get_level_dv": [vp, C.c_int, C.c_int, dp],
            "mldhip_get_omega": [vp, C.c_int, dp],
            "mldhip_get_tri_levels": [vp, C.c_int, C.c_int, ip, ip],
            "mldhip_get_dense_tail": [vp, ip],
            "mldhip_get_complexity": [vp, dp],
            "mldhip_dump": [vp, C.c_int, C.c_int, cp, cp, C.c_int],
            "mldhip_get_device_bytes": [vp, i64p],
            "mldhip_get_error": [vp, C.c_char_p, C.c_int],
            "mldhip_level_spmv": [vp, C.c_int, C.c_int, vp, vp],
            "mldhip_level_smooth": [vp, C.c_int, C.c_int, vp, C.c_int, vp, C.c_int, C.c_char],
            "mldhip_time_step": [vp, C.c_int, C.c_int, C.c_int, dp],
            "mldhip_get_op_format": [vp, C.c_int, C.c_int, ip, ip, ip],
            "mldhip_krylov_adjoint": [vp, cp, vp, vp, C.c_double, C.c_int, C.c_int, ip, dp],
            "mldhip_fv7_grad_theta": [C.c_int, vp, vp, vp, vp, vp, vp],
            "mldhip_apply_block": [vp, C.c_int, vp, C.c_int64, vp, C.c_int64],
            "mldhip_krylov_block": [vp, cp, C.c_int, vp, C.c_int64, vp, C.c_int64, C.c_double, C.c_int, C.c_int, ip,
                                    dp],
            "mldhip_group_create": [C.POINTER(vp), C.c_int, C.c_int],
            "mldhip_group_destroy": [vp],
            "mldhip_group_add": [vp, vp],
            "mldhip_group_build": [vp, C.c_int],
            "mldhip_group_info": [vp, ip, i64p, i64p],
            "mldhip_group_set_stream": [vp, vp],
            "mldhip_group_apply": [vp, C.c_int, vp, C.c_int64, vp, C.c_int64],
            "mldhip_group_krylov": [vp, cp, C.c_int, vp, C.c_int64, vp, C.c_int64, C.c_double, C.c_int, C.c_int, ip,
                                    dp, ip],
            "mldhip_group_time_apply": [vp, C.c_int, dp],
            "mldhip_group_level_info": [vp, C.c_int, i64p],
            "mldhip_group_get_error": [vp, C.c_char_p, C.c_int],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ptr(t):
    """Device pointer of a torch tensor (or an int address)."""
    if isinstance(t, int):
        return C.c_void_p(t)
    return C.c_void_p(t.data_ptr())


def _b(s):
    return None if s is None else s.encode()


def csr_arrays(a):
    """(n, irp int64, ja int32, val float64) from a scipy matrix or a tuple."""
    if isinstance(a, tuple):
        irp, ja, val = a
    else:
        m = a.tocsr()
        m.sort_indices()
        irp, ja, val = m.indptr, m.indices, m.data
    irp = np.ascontiguousarray(irp, dtype=np.int64)
    ja = np.ascontiguousarray(ja, dtype=np.int32)
    val = np.ascontiguousarray(val, dtype=np.float64)
    return len(irp) - 1, irp, ja, val


LEVEL_OPS = {"A": 0, "P": 1, "R": 2, "ND": 3, "L": 4, "U": 5, "ND2": 13, "L2": 14, "U2": 15}


class MldPrec:
    """type(mld_dprec_type) on the MI355X engine (one handle per object)."""

    def __init__(self, device: int = 0):
        self.h = C.c_void_p()
        rc = lib().mldhip_create(C.byref(self.h), device)
        if rc:
            raise MldError(rc, f"cannot create preconditioner on device {device}")
        self.device = device
        self.n = None

    # -- errors -----------------------------------------------------------
    def _check(self, rc, what):
        if rc:
            buf = C.create_string_buffer(512)
            lib().mldhip_get_error(self.h, buf, 512)
            raise MldError(rc, f"{what}: {buf.value.decode(errors='replace')}")

    def free(self):
        if self.h:
            lib().mldhip_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    # -- configuration ------------------------------------------------------
    def init(self, ptype: str):
        self._check(lib().mldhip_init(self.h, ptype.encode()), f"init({ptype})")
        return self

    def set(self, what: str, val, ilev: int | None = None, ilmax: int | None = None, pos: str | None = None):
        il = -1 if ilev is None else int(ilev)
        im = -1 if ilmax is None else int(ilmax)
        if isinstance(val, str):
            rc = lib().mldhip_setc(self.h, what.encode(), val.encode(), il, im, _b(pos))
        elif isinstance(val, (float, np.floating)):
            rc = lib().mldhip_setr(self.h, what.encode(), float(val), il, im, _b(pos))
        else:
            rc = lib().mldhip_seti(self.h, what.encode(), int(val), il, im, _b(pos))
        self._check(rc, f"set({what}={val!r})")
        return self

    def set_precision(self, kind: str):
        """'double' (FP64 hierarchy) or 'mixed' (FP32 hierarchy, FP64 vectors)."""
        self._check(lib().mldhip_set_precision(self.h, {"double": 0, "mixed": 1}[kind]), "set_precision")
        return self

    def set_format(self, fmt: str):
        self._check(lib().mldhip_set_format(self.h, {"auto": 0, "csr": 1, "ell": 2, "sell": 3, "sell256": 4, "sellu": 5, "dia": 6,
                                                          "sell128": 7, "seld": 8, "seld4096": 9,
                                                          "seld128": 10}[fmt]), "set_format")
        return self

    def set_setup(self, where: str):
        """Where hierarchy_build forms the level products: 'host', 'device'
        or 'auto' (the device for levels of 2^22+ rows when the handle has
        one)."""
        self._check(lib().mldhip_set_setup(self.h, {"auto": -1, "host": 0, "device": 1}[where]), "set_setup")
        return self

    # -- build ----------------------------------------------------------------
    def hierarchy_build(self, a):
        n, irp, ja, val = csr_arrays(a)
        self._keep = (irp, ja, val)
        rc = lib().mldhip_hierarchy_build(self.h, n, irp.ctypes.data_as(C.POINTER(C.c_int64)),
                                          ja.ctypes.data_as(C.POINTER(C.c_int)), _dp(val))
        self._check(rc, "hierarchy_build")
        self.n = n
        return self

    def smoothers_build(self):
        self._check(lib().mldhip_smoothers_build(self.h), "smoothers_build")
        return self

    def build(self, a):
        return self.hierarchy_build(a).smoothers_build()

    # -- a hierarchy built elsewhere ------------------------------------------------
    SOLVERS = {"NONE": 4, "DIAG": 5, "GS": 6, "ILU": 7, "BWGS": 14, "L1-DIAG": 101}

    def upload_level(self, lev, a, p=None, r=None):
        """Level lev (0-based) of an externally built hierarchy: A, and for
        lev >= 1 the prolongation P (n_{lev-1} x n) and restriction R."""
        n, irp, ja, val = csr_arrays(a)
        keep = [(irp, ja, val)]
        if lev > 0:
            _, pi, pj, pv = csr_arrays(p)
            _, ri, rj, rv = csr_arrays(r)
            nf = p.shape[0]
            keep += [(pi, pj, pv), (ri, rj, rv)]
        args = []
        for t in keep + [None] * (3 - len(keep)):
            if t is None:
                args += [None, None, None]
            else:
                args += [t[0].ctypes.data_as(C.POINTER(C.c_int64)), t[1].ctypes.data_as(C.POINTER(C.c_int)),
                         _dp(t[2])]
        self._check(lib().mldhip_upload_level(self.h, lev, n, *args[:3], nf if lev > 0 else 0, *args[3:]),
                    f"upload_level({lev})")
        if lev == 0:
            self.n = n
        return self

    def upload_smoother(self, lev, solver, dv=None, nd=None, l=None, u=None, which=0):
        """Smoother data of level lev: solver in SOLVERS, dv, ND, strict L, U."""
        def csr(m):
            if m is None:
                return [None, None, None], None
            _, i, j, v = csr_arrays(m)
            return [i.ctypes.data_as(C.POINTER(C.c_int64)), j.ctypes.data_as(C.POINTER(C.c_int)), _dp(v)], (i, j, v)
        a_nd, k1 = csr(nd)
        a_l, k2 = csr(l)
        a_u, k3 = csr(u)
        d = None if dv is None else np.ascontiguousarray(dv, dtype=np.float64)
        rc = lib().mldhip_upload_smoother(self.h, lev, which, self.SOLVERS[solver], *a_nd,
                                          None if d is None else _dp(d), *a_l, *a_u)
        self._check(rc, f"upload_smoother({lev})")
        return self

    def finalize_hierarchy(self, nlev):
        self._check(lib().mldhip_finalize_hierarchy(self.h, nlev), "finalize_hierarchy")
        return self

    def load_hierarchy(self, prefix, a, smoothers=False):
        """Hierarchy of a dump under prefix over the fine matrix a."""
        n, irp, ja, val = csr_arrays(a)
        self._keep = (irp, ja, val)
        rc = lib().mldhip_load_hierarchy(self.h, prefix.encode(), n, irp.ctypes.data_as(C.POINTER(C.c_int64)),
                                         ja.ctypes.data_as(C.POINTER(C.c_int)), _dp(val), int(bool(smoothers)))
        self._check(rc, "load_hierarchy")
        self.n = n
        return self

    # -- apply ----------------------------------------------------------------
    def apply(self, x, y=None, trans: str = "N"):
        """y = op(M^-1) x.  torch CUDA tensors (float64) stay on the device;
        numpy arrays go through the host entry point."""
        if isinstance(x, np.ndarray):
            x = np.ascontiguousarray(x, dtype=np.float64)
            y = np.zeros_like(x) if y is None else y
            self._check(lib().mldhip_apply_host(self.h, _dp(x), _dp(y), trans.encode()), "apply")
            return y
        import torch

        if y is None:
            y = torch.empty_like(x)
        self.use_torch_stream()
        self._check(lib().mldhip_apply(self.h, _ptr(x), _ptr(y), trans.encode()), "apply")
        return y

    def mlprec_aply(self, alpha, x, beta, y, trans: str = "N"):
        """y = alpha op(M^-1) x + beta y (mld_dmlprec_aply_vect); torch CUDA
        float64 tensors."""
        self.use_torch_stream()
        self._check(lib().mldhip_mlprec_aply(self.h, float(alpha), _ptr(x), float(beta), _ptr(y), trans.encode()),
                    "mlprec_aply")
        return y

    def apply_block(self, X, Y=None):
        """Y = M^-1 X for up to 32 columns at once (block V-cycle).  X: float64
        CUDA tensor of shape (nrhs, n), one column of the multi-vector per
        row (column-major storage, leading dimension n)."""
        import torch

        if X.dim() != 2 or not X.is_contiguous():
            raise MldError(-1, "apply_block: contiguous (nrhs, n) tensor expected")
        Y = torch.empty_like(X) if Y is None else Y
        self.use_torch_stream()
        k, n = X.shape
        self._check(lib().mldhip_apply_block(self.h, k, _ptr(X), n, _ptr(Y), n), "apply_block")
        return Y

    def krylov_block(self, method, B, X=None, eps=1e-8, itmax=1000, istop=2):
        """psb_krylov on every column of B (shape (nrhs, n)) at once.
        Returns (X, iters, errs) with per-column numpy arrays."""
        import torch

        k, n = B.shape
        X = torch.zeros_like(B) if X is None else X
        it = np.zeros(k, dtype=np.int32)
        er = np.zeros(k, dtype=np.float64)
        self.use_torch_stream()
        self._check(lib().mldhip_krylov_block(self.h, method.encode(), k, _ptr(B), n, _ptr(X), n, eps, itmax, istop,
                                              it.ctypes.data_as(C.POINTER(C.c_int)), _dp(er)), "krylov_block")
        return X, it, er

    def use_torch_stream(self):
        import torch

        s = torch.cuda.current_stream(self.device).cuda_stream
        lib().mldhip_set_stream(self.h, C.c_void_p(s))

    def synchronize(self):
        self._check(lib().mldhip_synchronize(self.h), "synchronize")

    # -- inspection --------------------------------------------------------------
    @property
    def nlev(self):
        v = C.c_int(0)
        self._check(lib().mldhip_get_nlev(self.h, C.byref(v)), "get_nlev")
        return v.value

    def level_info(self, lev):
        info = np.zeros(10, dtype=np.int64)
        self._check(lib().mldhip_get_level_info(self.h, lev, info.ctypes.data_as(C.POINTER(C.c_int64))),
                    "get_level_info")
        keys = ["n", "nnz_a", "nnz_p", "nnz_r", "nnz_nd", "naggr", "sweeps_pre", "sweeps_post", "solver",
                "solver_post"]
        return dict(zip(keys, (int(v) for v in info)))

    def level_csr(self, lev, what="A"):
        code = LEVEL_OPS[what]
        nr, nc, nz = C.c_int64(), C.c_int64(), C.c_int64()
        self._check(lib().mldhip_get_csr_size(self.h, lev, code, C.byref(nr), C.byref(nc), C.byref(nz)),
                    f"level_csr({lev},{what})")
        irp = np.zeros(nr.value + 1, dtype=np.int64)
        ja = np.zeros(nz.value, dtype=np.int32)
        val = np.zeros(nz.value)
        self._check(lib().mldhip_get_level_csr(self.h, lev, code, irp.ctypes.data_as(C.POINTER(C.c_int64)),
                                               ja.ctypes.data_as(C.POINTER(C.c_int)), _dp(val)),
                    "get_level_csr")
        import scipy.sparse as sp

        return sp.csr_matrix((val, ja, irp), shape=(nr.value, nc.value))

    def aggr(self, lev):
        n = self.level_info(lev - 1)["n"]
        out = np.zeros(n, dtype=np.int32)
        self._check(lib().mldhip_get_aggr(self.h, lev, out.ctypes.data_as(C.POINTER(C.c_int))), "get_aggr")
        return out

    def dv(self, lev, which=0):
        out = np.zeros(self.level_info(lev)["n"])
        self._check(lib().mldhip_get_level_dv(self.h, lev, which, _dp(out)), "get_level_dv")
        return out

    def dense_tail(self):
        """First level of the dense tail operator, or -1."""
        v = C.c_int()
        self._check(lib().mldhip_get_dense_tail(self.h, C.byref(v)), "get_dense_tail")
        return v.value

    def tri_levels(self, lev, which=0):
        """(level sets of the lower solve, of the upper solve) of a smoother."""
        a, b = C.c_int(), C.c_int()
        self._check(lib().mldhip_get_tri_levels(self.h, lev, which, C.byref(a), C.byref(b)), "get_tri_levels")
        return a.value, b.value

    def omega(self, lev):
        v = C.c_double()
        self._check(lib().mldhip_get_omega(self.h, lev, C.byref(v)), "get_omega")
        return v.value

    def complexity(self):
        v = C.c_double()
        self._check(lib().mldhip_get_complexity(self.h, C.byref(v)), "get_complexity")
        return v.value

    def dump(self, prefix=None, istart=None, iend=None, head=None, ac=False, rp=False, smoother=False,
             solver=False):
        """prec%dump (mlprec/mld_d_prec_type.f90:698-729): MatrixMarket files
        <prefix>_p000_lLLL_{ac,r,p}.mtx and <prefix>_sm[2a]_p000_lLLL_
        {nd,diag,lower,upper}.mtx for levels istart..iend (1-based)."""
        flags = (1 if ac else 0) | (2 if rp else 0) | (4 if smoother else 0) | (8 if solver else 0)
        self._check(lib().mldhip_dump(self.h, istart or 0, iend or 0, (prefix or "").encode(),
                                      (head or "").encode(), flags), "dump")
        return self

    def device_bytes(self):
        v = C.c_int64()
        self._check(lib().mldhip_get_device_bytes(self.h, C.byref(v)), "get_device_bytes")
        return v.value

    def descr(self):
        """Text description of the hierarchy (prec%descr analogue)."""
        lines = [f"Multilevel preconditioner: {self.nlev} levels, operator complexity {self.complexity():.4f}"]
        for l in range(self.nlev):
            i = self.level_info(l)
            lines.append(f"  level {l + 1}: n={i['n']} nnz(A)={i['nnz_a']} nnz(P)={i['nnz_p']} "
                         f"sweeps={i['sweeps_pre']}/{i['sweeps_post']} solver={i['solver']}")
        return "\n".join(lines)

    # -- single kernels ---------------------------------------------------------
    def level_spmv(self, lev, what, x, y):
        self.use_torch_stream()
        self._check(lib().mldhip_level_spmv(self.h, lev, LEVEL_OPS[what], _ptr(x), _ptr(y)), "level_spmv")
        return y

    STEP_KINDS = {"residual": 0, "pre": 1, "restrict": 2, "prolong": 3, "post": 4, "krylov_spmv": 5,
                  "residual_in_cycle": 6}

    def time_step(self, lev, kind, reps=20):
        """Average ms of one cycle step (HIP events on the handle's stream)."""
        self.use_torch_stream()
        v = C.c_double()
        self._check(lib().mldhip_time_step(self.h, lev, self.STEP_KINDS[kind], reps, C.byref(v)), "time_step")
        return v.value

    def op_format(self, lev, what):
        f, l, w = C.c_int(), C.c_int(), C.c_int()
        self._check(lib().mldhip_get_op_format(self.h, lev, LEVEL_OPS[what], C.byref(f), C.byref(l), C.byref(w)),
                    "get_op_format")
        return {"fmt": ("CSR", "ELL", "SELL", "DIA", "SELD")[f.value], "lanes": l.value, "width": w.value}

    def level_smooth(self, lev, x, y, beta=0, sweeps=1, trans="N", which=0):
        self.use_torch_stream()
        self._check(lib().mldhip_level_smooth(self.h, lev, which, _ptr(x), beta, _ptr(y), sweeps, trans.encode()),
                    "level_smooth")
        return y


def psb_krylov(method, a, prec: MldPrec, b, x=None, eps=1e-6, itmax=500, istop=2):
    """psb_krylov(method, a, prec, b, x, eps, desc, info, itmax, iter, err,
    istop) with the matrix `a` given to prec.hierarchy_build (the device
    keeps its copy).  Returns (x, iter, err)."""
    it = C.c_int(0)
    err = C.c_double(0.0)
    if isinstance(b, np.ndarray):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b) if x is None else np.ascontiguousarray(x, dtype=np.float64).copy()
        rc = lib().mldhip_krylov_host(prec.h, method.encode(), _dp(b), _dp(x), eps, itmax, istop, C.byref(it),
                                      C.byref(err))
    else:
        import torch

        x = torch.zeros_like(b) if x is None else x
        prec.use_torch_stream()
        rc = lib().mldhip_krylov(prec.h, method.encode(), _ptr(b), _ptr(x), eps, itmax, istop, C.byref(it),
                                 C.byref(err))
    prec._check(rc, f"krylov({method})")
    return x, it.value, err.value


def psb_krylov_adjoint(method, prec: MldPrec, b, x=None, eps=1e-6, itmax=500, istop=2):
    """Adjoint solve A^T x = b with the hierarchy of `prec` applied
    transposed (prec%apply(..., trans='T')).  b, x: float64 CUDA tensors.
    Returns (x, iter, err)."""
    import torch

    it = C.c_int(0)
    err = C.c_double(0.0)
    x = torch.zeros_like(b) if x is None else x
    prec.use_torch_stream()
    rc = lib().mldhip_krylov_adjoint(prec.h, method.encode(), _ptr(b), _ptr(x), eps, itmax, istop, C.byref(it),
                                     C.byref(err))
    prec._check(rc, f"krylov_adjoint({method})")
    return x, it.value, err.value


def fv7_grad_theta(idim, theta, x, lam, grad=None, work=True):
    """grad[c] = -lam^T (dA/dtheta_c) x for the 7-point finite-volume
    diffusion operator with k = exp(theta) (device kernel, torch's current
    stream).  All tensors float64 on the GPU, length idim^3.  work: a scratch
    tensor for exp(theta), True to allocate one, or None for the variant that
    forms the exponentials in place."""
    import torch

    for t in (theta, x, lam):
        if t.dtype != torch.float64 or not t.is_cuda or t.numel() != idim ** 3 or not t.is_contiguous():
            raise MldError(-1, "fv7_grad_theta: float64 contiguous CUDA tensors of length idim^3 expected")
    grad = torch.empty_like(theta) if grad is None else grad
    if work is True:
        work = torch.empty_like(theta)
    s = torch.cuda.current_stream(theta.device).cuda_stream
    rc = lib().mldhip_fv7_grad_theta(idim, _ptr(theta), _ptr(x), _ptr(lam), _ptr(grad),
                                     None if work is None else _ptr(work), C.c_void_p(s))
    if rc:
        raise MldError(rc, "fv7_grad_theta")
    return grad


class MldGroup:
    """Independent systems solved together (grouped V-cycle + one hipGraph
    for the Krylov iterations of all systems).  nb = 1: one column per
    system; nb = 32: up to 32 columns each."""

    def __init__(self, device: int = 0, nb: int = 1):
        self.h = C.c_void_p()
        self.device = device
        self._sys = []
        rc = lib().mldhip_group_create(C.byref(self.h), device, nb)
        if rc:
            raise MldError(rc, "group_create")

    def _check(self, rc, what):
        if rc:
            buf = C.create_string_buffer(512)
            lib().mldhip_group_get_error(self.h, buf, 512)
            raise MldError(rc, f"{what}: {buf.value.decode(errors='replace')}")

    def __del__(self):
        try:
            if self.h:
                lib().mldhip_group_destroy(self.h)
                self.h = C.c_void_p()
        except Exception:
            pass

    def add(self, prec: MldPrec):
        self._check(lib().mldhip_group_add(self.h, prec.h), "group_add")
        self._sys.append(prec)
        return self

    def build(self, precision: str = "double"):
        self._check(lib().mldhip_group_build(self.h, {"double": 0, "mixed": 1}[precision]), "group_build")
        return self

    def info(self):
        ns = C.c_int(0)
        nr = C.c_int64(0)
        self._check(lib().mldhip_group_info(self.h, C.byref(ns), C.byref(nr), None), "group_info")
        off = np.zeros(ns.value + 1, dtype=np.int64)
        self._check(lib().mldhip_group_info(self.h, None, None, off.ctypes.data_as(C.POINTER(C.c_int64))),
                    "group_info")
        return ns.value, nr.value, off

    def _stream(self):
        import torch

        s = torch.cuda.current_stream(self.device).cuda_stream
        lib().mldhip_group_set_stream(self.h, C.c_void_p(s))

    def apply(self, X, Y=None):
        """X: float64 CUDA tensor (nrows,) or (nrhs, nrows)."""
        import torch

        X2 = X if X.dim() == 2 else X.view(1, -1)
        Y = torch.empty_like(X) if Y is None else Y
        Y2 = Y if Y.dim() == 2 else Y.view(1, -1)
        k, n = X2.shape
        self._stream()
        self._check(lib().mldhip_group_apply(self.h, k, _ptr(X2), n, _ptr(Y2), n), "group_apply")
        return Y

    def krylov(self, method, B, X=None, eps=1e-8, itmax=1000, istop=2):
        """Returns (X, iters, errs, graph_launches); iters/errs of shape
        (nrhs, nsys)."""
        import torch

        B2 = B if B.dim() == 2 else B.view(1, -1)
        X = torch.zeros_like(B) if X is None else X
        X2 = X if X.dim() == 2 else X.view(1, -1)
        k, n = B2.shape
        ns = len(self._sys)
        it = np.zeros(k * ns, dtype=np.int32)
        er = np.zeros(k * ns, dtype=np.float64)
        nl = C.c_int(0)
        self._stream()
        self._check(lib().mldhip_group_krylov(self.h, method.encode(), k, _ptr(B2), n, _ptr(X2), n, eps, itmax,
                                              istop, it.ctypes.data_as(C.POINTER(C.c_int)), _dp(er), C.byref(nl)),
                    "group_krylov")
        return X, it.reshape(k, ns), er.reshape(k, ns), nl.value

    def time_apply(self, reps=10):
        ms = C.c_double(0.0)
        self._check(lib().mldhip_group_time_apply(self.h, reps, C.byref(ms)), "group_time_apply")
        return ms.value

    def level_info(self, lev):
        info = np.zeros(10, dtype=np.int64)
        self._check(lib().mldhip_group_level_info(self.h, lev, info.ctypes.data_as(C.POINTER(C.c_int64))),
                    "group_level_info")
        keys = ["n", "ni", "nnz_a", "nnz_p", "nnz_r", "nnz_nd", "nnz_coarse", "sweeps_pre", "sweeps_post", "storage"]
        return dict(zip(keys, (int(v) for v in info)))
