"""pastix_amd -- PaStiX's C API served by a single-MI355X numerical engine.

This package loads the in-tree C-ABI library ``libpastix_amd.so`` (host C +
HIP kernels for gfx950) and mirrors the reference's Python wrapper
(``wrappers/python/examples/pypastix/pastix.py``): ``initParam``, ``init``,
``task_analyze``, ``task_numfact``, ``task_solve``, ``task_refine``,
``subtask_*`` and ``finalize`` keep their names, argument meaning and error
behaviour (a ``PastixError`` carrying the pastix_error_t code).

There is no host fallback: the numerical steps raise when no GPU is usable.
"""
import ctypes
import os
import re

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_ROOT = os.path.dirname(_HERE)
LIB_PATH = os.path.join(_HERE, "libpastix_amd.so")

if not os.path.exists(LIB_PATH):
    raise ImportError("pastix_amd: %s is missing; run `python -m pastix_amd.build`" % LIB_PATH)
lib = ctypes.CDLL(LIB_PATH)

pastix_int = ctypes.c_int64
_p = ctypes.c_void_p


def _parse_enums(header):
    """Names and values of the enums of a C header (sequential or explicit)."""
    out = {}
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    for body in re.findall(r"typedef\s+enum\s+\w*\s*\{(.*?)\}", text, flags=re.S):
        val = -1
        for item in body.split(","):
            item = item.strip()
            if not item:
                continue
            if "=" in item:
                name, v = [s.strip() for s in item.split("=")]
                val = out[v] if v in out else int(v, 0)
            else:
                name = item
                val += 1
            out[name] = val
    return out


ENUMS = _parse_enums(os.path.join(_ROOT, "include", "pastix", "api.h"))
ENUMS.update(_parse_enums(os.path.join(_ROOT, "include", "spm.h")))
globals().update(ENUMS)
PastixDouble = ENUMS["SpmDouble"]
PastixFloat = ENUMS["SpmFloat"]
PastixComplex64 = ENUMS["SpmComplex64"]
PastixComplex32 = ENUMS["SpmComplex32"]
PastixGeneral = ENUMS["SpmGeneral"]
PastixSymmetric = ENUMS["SpmSymmetric"]
PastixHermitian = ENUMS["SpmHermitian"]


class PastixError(RuntimeError):
    def __init__(self, func, code):
        super().__init__("%s returned pastix_error_t %d" % (func, code))
        self.code = code


class spmatrix_t(ctypes.Structure):
    _fields_ = [("mtxtype", ctypes.c_int), ("flttype", ctypes.c_int), ("fmttype", ctypes.c_int),
                ("baseval", pastix_int), ("gN", pastix_int), ("n", pastix_int),
                ("gnnz", pastix_int), ("nnz", pastix_int), ("gNexp", pastix_int),
                ("nexp", pastix_int), ("gnnzexp", pastix_int), ("nnzexp", pastix_int),
                ("dof", pastix_int), ("dofs", _p), ("layout", ctypes.c_int),
                ("colptr", _p), ("rowptr", _p), ("loc2glob", _p), ("values", _p),
                ("clustnum", ctypes.c_int), ("clustnbr", ctypes.c_int)]


class pastix_order_t(ctypes.Structure):
    _fields_ = [("baseval", pastix_int), ("vertnbr", pastix_int), ("cblknbr", pastix_int),
                ("permtab", ctypes.POINTER(pastix_int)), ("peritab", ctypes.POINTER(pastix_int)),
                ("rangtab", ctypes.POINTER(pastix_int)), ("treetab", ctypes.POINTER(pastix_int)),
                ("selevtx", _p), ("sndenbr", pastix_int), ("sndetab", _p), ("peritab_exp", _p)]


def _sig(name, restype, *args):
    f = getattr(lib, name)
    f.restype = restype
    f.argtypes = list(args)
    return f


_I = ctypes.c_int
_sig("pastixInitParam", None, _p, _p)
_sig("pastixInit", None, ctypes.POINTER(_p), ctypes.c_size_t, _p, _p)
_sig("pastixFinalize", None, ctypes.POINTER(_p))
for _n in ("pastix_task_analyze", "pastix_task_numfact", "pastix_subtask_spm2bcsc"):
    _sig(_n, _I, _p, ctypes.POINTER(spmatrix_t))
_sig("pastix_subtask_order", _I, _p, ctypes.POINTER(spmatrix_t), ctypes.POINTER(pastix_order_t))
for _n in ("pastix_subtask_symbfact", "pastix_subtask_reordering", "pastix_subtask_blend",
           "pastix_subtask_bcsc2ctab", "pastix_subtask_sopalin"):
    _sig(_n, _I, _p)
_sig("pastix_task_solve", _I, _p, pastix_int, pastix_int, _p, pastix_int)
_sig("pastix_task_refine", _I, _p, pastix_int, pastix_int, _p, pastix_int, _p, pastix_int)
_sig("pastix_task_solve_and_refine", _I, _p, pastix_int, pastix_int, _p, pastix_int, _p, pastix_int)
_sig("pastixOrderGrid", _I, ctypes.POINTER(ctypes.POINTER(pastix_order_t)), pastix_int, pastix_int, pastix_int)
_sig("pastixOrderExit", None, ctypes.POINTER(pastix_order_t))
_sig("pastixOrderCheck", _I, ctypes.POINTER(pastix_order_t))
_sig("pastixOrderExpand", None, ctypes.POINTER(pastix_order_t), ctypes.POINTER(spmatrix_t))
_sig("pastixOrderSave", _I, _p, ctypes.POINTER(pastix_order_t))
_sig("pastixOrderLoad", _I, _p, ctypes.POINTER(pastix_order_t))
_sig("pastixOrderAllocId", _I, ctypes.POINTER(pastix_order_t), pastix_int)
_sig("pastixOrderGet", ctypes.POINTER(pastix_order_t), _p)
_sig("pastixExpand", None, _p, ctypes.POINTER(spmatrix_t))
_sig("pastixGetOptions", None, _I, _p, _p, _p, _p, _p, _p, _p)
_sig("pastix_amd_symbol", _p, _p)
_sig("pastixSymbolSave", _I, _p, _p)
_sig("pastixSymbolLoad", _I, _p, _p)
_sig("pastixSymbolExit", None, _p)
_sig("pastix_amd_parse_iparm", _I, ctypes.c_char_p)
_sig("pastix_amd_parse_dparm", _I, ctypes.c_char_p)
_sig("pastix_amd_parse_enum", _I, ctypes.c_char_p)
_sig("spmInit", None, ctypes.POINTER(spmatrix_t))
_sig("spmExit", None, ctypes.POINTER(spmatrix_t))
_sig("spmNorm", ctypes.c_double, _I, ctypes.POINTER(spmatrix_t))
_sig("spmMatMat", _I, _I, pastix_int, ctypes.c_double, ctypes.POINTER(spmatrix_t), _p, pastix_int,
     ctypes.c_double, _p, pastix_int)
_sig("spmGenRHSSeed", _I, _I, pastix_int, ctypes.POINTER(spmatrix_t), ctypes.c_ulonglong, _p, pastix_int,
     _p, pastix_int)
_sig("spmBackwardError", ctypes.c_double, pastix_int, ctypes.POINTER(spmatrix_t), _p, pastix_int, _p,
     pastix_int, _p)
_sig("spmCheckAxb", _I, ctypes.c_double, pastix_int, ctypes.POINTER(spmatrix_t), _p, pastix_int, _p, pastix_int,
     _p, pastix_int)
_sig("spmForwardError", ctypes.c_double, pastix_int, ctypes.POINTER(spmatrix_t), _p, pastix_int, _p, pastix_int)
_sig("spm_plrnt", None, pastix_int, ctypes.c_ulonglong, ctypes.c_ulonglong, _p)
_sig("spmReadDriver", _I, _I, ctypes.c_char_p, ctypes.POINTER(spmatrix_t))
_sig("spmGenLaplacian7", _I, ctypes.POINTER(spmatrix_t), _I, pastix_int, pastix_int, pastix_int,
     ctypes.c_double, ctypes.c_double)
_sig("spmGenConvDiff27", _I, ctypes.POINTER(spmatrix_t), pastix_int, pastix_int, pastix_int,
     ctypes.c_double, ctypes.c_double, ctypes.c_ulonglong)
_sig("spmGenHelmholtz7", _I, ctypes.POINTER(spmatrix_t), pastix_int, pastix_int, pastix_int,
     ctypes.c_double, ctypes.c_double)
_sig("spmGenVLaplacian7", _I, ctypes.POINTER(spmatrix_t), pastix_int, pastix_int, pastix_int, pastix_int,
     ctypes.c_double, ctypes.c_ulonglong)
_PI = ctypes.POINTER(pastix_int)
_sig("pastix_amd_solver_sizes", _I, _p, _PI, _PI, _PI, _PI)
_sig("pastix_amd_solver_export", _I, _p, _p, _p, _p)
_sig("pastix_amd_order_sizes", _I, _p, _PI, _PI)
_sig("pastix_amd_order_export", _I, _p, _p, _p, _p, _p)
_sig("pastix_amd_symbol_sizes", _I, _p, _PI, _PI)
_sig("pastix_amd_symbol_export", _I, _p, _p, _p)
_sig("pastix_amd_bcsc_sizes", _I, _p, _PI, _PI)
_sig("pastix_amd_bcsc_export", _I, _p, _p, _p, _p)
_sig("pastix_amd_get_coeftab", _I, _p, _I, _p, ctypes.c_size_t)
_sig("pastix_amd_get_stats", _I, _p, _p, _I)
_sig("pastix_amd_set_option", _I, _p, ctypes.c_char_p, ctypes.c_double)
_sig("pastix_amd_device_handle", _I, _p, ctypes.POINTER(_p))
_sig("pastix_amd_solve_dev", _I, _p, pastix_int, _p, pastix_int)
_sig("pastix_amd_refine_dev", _I, _p, pastix_int, _p, _p, _PI)
_sig("pastixRhsInit", _I, ctypes.POINTER(_p))
_sig("pastixRhsFinalize", _I, _p)
_sig("pastixRhsDoubletoSingle", _I, _p, _p)
_sig("pastixRhsSingleToDouble", _I, _p, _p)
_sig("pastix_subtask_applyorder", _I, _p, _I, pastix_int, pastix_int, _p, pastix_int, _p)
_sig("pastix_subtask_trsm", _I, _p, _I, _I, _I, _I, _p)
_sig("pastix_subtask_diag", _I, _p, _p)
_sig("pastix_subtask_solve", _I, _p, _p)
_sig("pastix_subtask_solve_adv", _I, _p, _I, _p)
_sig("pastix_subtask_refine", _I, _p, _p, _p)
_sig("pastixGetDiag", _I, _p, _p, pastix_int)
_sig("pastixSetSchurUnknownList", None, _p, pastix_int, _p)
_sig("pastixGetSchur", _I, _p, _p, pastix_int)
_sig("pastixRhsSchurGet", _I, _p, pastix_int, pastix_int, _p, _p, pastix_int)
_sig("pastixRhsSchurSet", _I, _p, pastix_int, pastix_int, _p, pastix_int, _p)
_sig("pxd_device_count", _I)
_sig("pxd_spmv", _I, _p, ctypes.c_double, _p, ctypes.c_double, _p, pastix_int, pastix_int)
_sig("pxd_malloc", _I, _p, ctypes.c_size_t, ctypes.POINTER(_p))
_sig("pxd_free", _I, _p, _p)
_sig("pxd_h2d", _I, _p, _p, _p, ctypes.c_size_t)
_sig("pxd_d2h", _I, _p, _p, _p, ctypes.c_size_t)
_sig("pxd_sync", _I, _p)
_sig("pxd_vec_copy", _I, _p, pastix_int, pastix_int, _p, _p)


def _check(name, rc):
    if rc != 0:
        raise PastixError(name, rc)


def _ptr(a):
    return a.ctypes.data_as(_p)


# ---------------------------------------------------------------------------
# SPM helper: an spmatrix_t whose arrays live in numpy memory or in C memory
# allocated by the generators (freed with spmExit).

class Spm:
    """Owns one spmatrix_t (C struct) and keeps referenced numpy arrays alive."""

    def __init__(self):
        self.s = spmatrix_t()
        lib.spmInit(ctypes.byref(self.s))
        self._owned_by_c = False
        self._keep = []

    @classmethod
    def laplacian7(cls, nx, ny, nz, diag=6.0, offd=1.0, complex_=False):
        m = cls()
        _check("spmGenLaplacian7", lib.spmGenLaplacian7(ctypes.byref(m.s),
               ENUMS["SpmComplex64"] if complex_ else ENUMS["SpmDouble"], nx, ny, nz, diag, offd))
        m._owned_by_c = True
        return m

    @classmethod
    def convdiff27(cls, nx, ny, nz, diffusion=1.0, vmax=0.5, seed=0):
        m = cls()
        _check("spmGenConvDiff27", lib.spmGenConvDiff27(ctypes.byref(m.s), nx, ny, nz, diffusion, vmax, seed))
        m._owned_by_c = True
        return m

    @classmethod
    def helmholtz7(cls, nx, ny, nz, kr=0.5, ki=0.1):
        m = cls()
        _check("spmGenHelmholtz7", lib.spmGenHelmholtz7(ctypes.byref(m.s), nx, ny, nz, kr, ki))
        m._owned_by_c = True
        return m

    @classmethod
    def vlaplacian7(cls, nx, ny, nz, dof=3, sigma=1.0, seed=0):
        m = cls()
        _check("spmGenVLaplacian7", lib.spmGenVLaplacian7(ctypes.byref(m.s), nx, ny, nz, dof, sigma, seed))
        m._owned_by_c = True
        return m

    @classmethod
    def read(cls, driver, filename):
        """spmReadDriver: driver is an spm_driver_t value (SpmDriverHB, SpmDriverMM, ...)."""
        m = cls()
        _check("spmReadDriver", lib.spmReadDriver(driver, str(filename).encode(), ctypes.byref(m.s)))
        m._owned_by_c = True
        return m

    @classmethod
    def from_csc(cls, n, colptr, rowptr, values, mtxtype=None, baseval=0):
        m = cls()
        colptr = np.ascontiguousarray(colptr, dtype=np.int64)
        rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
        values = np.ascontiguousarray(values)
        if values.dtype == np.complex128:
            m.s.flttype = ENUMS["SpmComplex64"]
        else:
            values = values.astype(np.float64)
            m.s.flttype = ENUMS["SpmDouble"]
        m.s.mtxtype = ENUMS["SpmGeneral"] if mtxtype is None else mtxtype
        m.s.fmttype = ENUMS["SpmCSC"]
        m.s.baseval = baseval
        m.s.n = m.s.gN = m.s.nexp = m.s.gNexp = n
        m.s.nnz = m.s.gnnz = m.s.nnzexp = m.s.gnnzexp = len(rowptr)
        m.s.dof = 1
        m.s.colptr = colptr.ctypes.data
        m.s.rowptr = rowptr.ctypes.data
        m.s.values = values.ctypes.data
        m._keep = [colptr, rowptr, values]
        return m

    @property
    def n(self):
        return self.s.n

    @property
    def is_complex(self):
        return self.s.flttype == ENUMS["SpmComplex64"]

    @property
    def dtype(self):
        return np.complex128 if self.is_complex else np.float64

    def arrays(self):
        """(colptr, rowptr, values) numpy copies of the stored CSC."""
        n, nnz = self.s.n, self.s.nnz
        cp = np.ctypeslib.as_array(ctypes.cast(self.s.colptr, _PI), shape=(n + 1,)).copy()
        rp = np.ctypeslib.as_array(ctypes.cast(self.s.rowptr, _PI), shape=(nnz,)).copy()
        if self.is_complex:
            v = np.ctypeslib.as_array(ctypes.cast(self.s.values, ctypes.POINTER(ctypes.c_double)),
                                      shape=(2 * nnz,)).copy().view(np.complex128)
        else:
            v = np.ctypeslib.as_array(ctypes.cast(self.s.values, ctypes.POINTER(ctypes.c_double)),
                                      shape=(nnz,)).copy()
        return cp, rp, v

    def norm(self, ntype="fro"):
        code = {"fro": ENUMS["SpmFrobeniusNorm"], "inf": ENUMS["SpmInfNorm"],
                "one": ENUMS["SpmOneNorm"], "max": ENUMS["SpmMaxNorm"]}[ntype]
        return lib.spmNorm(code, ctypes.byref(self.s))

    def matmat(self, x, alpha=1.0, beta=0.0, y=None):
        x = np.asfortranarray(x, dtype=self.dtype)
        nrhs = 1 if x.ndim == 1 else x.shape[1]
        y = np.zeros_like(x) if y is None else np.array(y, dtype=self.dtype, order="F", copy=True)
        _check("spmMatMat", lib.spmMatMat(ENUMS["PastixNoTrans"], nrhs, alpha, ctypes.byref(self.s),
                                          _ptr(x), self.n, beta, _ptr(y), self.n))
        return y

    def gen_rhs(self, nrhs=1, seed=0):
        """x uniform in [-1,1) from the plrnt generator (column j: seed+j), b = A x."""
        x = np.zeros((self.n, nrhs), dtype=self.dtype, order="F")
        b = np.zeros((self.n, nrhs), dtype=self.dtype, order="F")
        _check("spmGenRHS", lib.spmGenRHSSeed(ENUMS["SpmRhsRndX"], nrhs, ctypes.byref(self.s), seed,
                                              _ptr(x), self.n, _ptr(b), self.n))
        return x, b

    def backward_error(self, b, x):
        b = np.asfortranarray(b, dtype=self.dtype)
        x = np.asfortranarray(x, dtype=self.dtype)
        nrhs = 1 if b.ndim == 1 else b.shape[1]
        each = np.zeros(nrhs)
        worst = lib.spmBackwardError(nrhs, ctypes.byref(self.s), _ptr(b), self.n, _ptr(x), self.n,
                                     _ptr(each))
        return worst, each

    def __del__(self):
        try:
            if self._owned_by_c:
                lib.spmExit(ctypes.byref(self.s))
        except Exception:
            pass


def check_axb(spm, eps, x0, b, x):
    """spmCheckAxb (0: the solution passes); x0 may be None."""
    b = np.asfortranarray(b, dtype=spm.dtype)
    x = np.asfortranarray(x, dtype=spm.dtype)
    nrhs = 1 if b.ndim == 1 else b.shape[1]
    x0p = None
    if x0 is not None:
        x0 = np.asfortranarray(x0, dtype=spm.dtype)
        x0p = _ptr(x0)
    return lib.spmCheckAxb(eps, nrhs, ctypes.byref(spm.s), x0p, spm.n, _ptr(b), spm.n, _ptr(x), spm.n)


def forward_error(x0, x, spm=None):
    """max over columns of ||x - x0||_inf / ||x0||_inf (spmForwardError when an spm is given)."""
    x0 = np.asfortranarray(x0)
    x = np.asfortranarray(x, dtype=x0.dtype)
    if x0.ndim == 1:
        x0, x = x0[:, None], x[:, None]
    if spm is not None:
        return lib.spmForwardError(x0.shape[1], ctypes.byref(spm.s), _ptr(x0), spm.n, _ptr(x), spm.n)
    return float(max(np.abs(x[:, c] - x0[:, c]).max() / np.abs(x0[:, c]).max() for c in range(x0.shape[1])))


def plrnt(n, seed=0, shift=0):
    out = np.zeros(n)
    lib.spm_plrnt(n, seed, shift, _ptr(out))
    return out


# ---------------------------------------------------------------------------
# Ordering helpers

class Order:
    def __init__(self):
        self.ptr = ctypes.pointer(pastix_order_t())

    @classmethod
    def grid(cls, nx, ny, nz):
        o = cls()
        _check("pastixOrderGrid", lib.pastixOrderGrid(ctypes.byref(o.ptr), nx, ny, nz))
        return o

    def expand(self, dof):
        """Expand a node ordering to dof unknowns per node (pastixOrderExpand)."""
        stub = spmatrix_t()
        stub.dof = dof
        stub.baseval = 0
        lib.pastixOrderExpand(self.ptr, ctypes.byref(stub))
        return self

    def arrays(self):
        o = self.ptr.contents
        n, c = o.vertnbr, o.cblknbr
        return (np.ctypeslib.as_array(o.permtab, shape=(n,)).copy(),
                np.ctypeslib.as_array(o.peritab, shape=(n,)).copy(),
                np.ctypeslib.as_array(o.rangtab, shape=(c + 1,)).copy(),
                np.ctypeslib.as_array(o.treetab, shape=(c,)).copy())

    def __del__(self):
        try:
            lib.pastixOrderExit(self.ptr)
        except Exception:
            pass


# ---------------------------------------------------------------------------
# pypastix-compatible functions

def initParam():
    iparm = np.zeros(ENUMS["IPARM_SIZE"], dtype=np.int64)
    dparm = np.zeros(ENUMS["DPARM_SIZE"], dtype=np.float64)
    lib.pastixInitParam(_ptr(iparm), _ptr(dparm))
    return iparm, dparm


# pastix_data keeps pointers to the caller's iparm/dparm arrays (the C API
# references them, reference common/api.c:767-768): hold them until finalize.
_LIVE_PARAMS = {}


def init(iparm, dparm, comm=0):
    if iparm.dtype != np.int64 or dparm.dtype != np.float64 or not iparm.flags.c_contiguous:
        raise TypeError("iparm must be a contiguous int64 array and dparm a float64 array")
    pd = _p()
    lib.pastixInit(ctypes.byref(pd), comm, _ptr(iparm), _ptr(dparm))
    _LIVE_PARAMS[pd.value] = (iparm, dparm)
    return pd


def finalize(pastix_data, iparm=None, dparm=None):
    key = pastix_data.value
    lib.pastixFinalize(ctypes.byref(pastix_data))
    _LIVE_PARAMS.pop(key, None)


def task_analyze(pastix_data, spm):
    _check("pastix_task_analyze", lib.pastix_task_analyze(pastix_data, ctypes.byref(spm.s)))


def subtask_order(pastix_data, spm, order=None):
    _check("pastix_subtask_order", lib.pastix_subtask_order(
        pastix_data, ctypes.byref(spm.s), order.ptr if order is not None else None))


def subtask_symbfact(pastix_data):
    _check("pastix_subtask_symbfact", lib.pastix_subtask_symbfact(pastix_data))


def subtask_reordering(pastix_data):
    _check("pastix_subtask_reordering", lib.pastix_subtask_reordering(pastix_data))


def subtask_blend(pastix_data):
    _check("pastix_subtask_blend", lib.pastix_subtask_blend(pastix_data))


def analyze_grid(pastix_data, spm, nx, ny, nz, dof=1):
    """Analysis with the grid nested dissection given as a personal ordering,
    the sequence of the reference's example/ordering_grid.c:107-123.  With
    dof > 1 the matrix holds dof unknowns per grid node (numbered node-major)
    and the node ordering is expanded to them (pastixOrderExpand)."""
    order = Order.grid(nx, ny, nz)
    if dof > 1:
        order.expand(dof)
    subtask_order(pastix_data, spm, order)
    subtask_symbfact(pastix_data)
    subtask_reordering(pastix_data)
    subtask_blend(pastix_data)
    return order


def task_numfact(pastix_data, spm):
    _check("pastix_task_numfact", lib.pastix_task_numfact(pastix_data, ctypes.byref(spm.s)))


def _rhs_array(name, a, spm, nrhs):
    """Check a right-hand-side array handed to the C API: an ndarray of the
    matrix's dtype, n rows, column-major (leading dimension n) and at least
    nrhs columns, so that the C side reads and writes exactly its memory."""
    if not isinstance(a, np.ndarray):
        raise ValueError("%s: expected a numpy.ndarray" % name)
    if a.dtype != spm.dtype:
        raise ValueError("%s: dtype %s does not match the matrix (%s)" % (name, a.dtype, np.dtype(spm.dtype)))
    if a.ndim not in (1, 2) or a.shape[0] != spm.n:
        raise ValueError("%s: expected %d rows" % (name, spm.n))
    if a.ndim == 2 and not a.flags.f_contiguous:
        raise ValueError("%s: a 2-D right-hand side must be Fortran-ordered (column-major)" % name)
    if a.ndim == 1 and not a.flags.c_contiguous:
        raise ValueError("%s: a 1-D right-hand side must be contiguous" % name)
    cols = 1 if a.ndim == 1 else a.shape[1]
    if nrhs < 0:
        nrhs = cols
    if nrhs > cols:
        raise ValueError("%s: %d right-hand sides asked for, %d given" % (name, nrhs, cols))
    return nrhs


def task_solve(pastix_data, spm, x, nrhs=-1):
    nrhs = _rhs_array("task_solve: x", x, spm, nrhs)
    _check("pastix_task_solve", lib.pastix_task_solve(pastix_data, spm.n, nrhs, _ptr(x), spm.n))


def task_refine(pastix_data, spm, b, x, nrhs=-1):
    nrhs = _rhs_array("task_refine: b", b, spm, nrhs)
    _rhs_array("task_refine: x", x, spm, nrhs)
    _check("pastix_task_refine", lib.pastix_task_refine(pastix_data, spm.n, nrhs, _ptr(b), spm.n,
                                                         _ptr(x), spm.n))


# ---------------------------------------------------------------------------
# Right-hand-side descriptors and the solve subtasks (pypastix's rhsInit,
# subtask_applyorder, subtask_trsm, subtask_diag, ...)

class pastix_rhs_s(ctypes.Structure):
    _fields_ = [("allocated", ctypes.c_int8), ("flttype", ctypes.c_int), ("m", pastix_int),
                ("n", pastix_int), ("ld", pastix_int), ("b", _p), ("cblkb", _p), ("rhscomm", _p),
                ("Ploc2Pglob", _p)]


def rhsInit():
    rhs = _p()
    _check("pastixRhsInit", lib.pastixRhsInit(ctypes.byref(rhs)))
    return rhs


def rhsFinalize(rhs):
    _check("pastixRhsFinalize", lib.pastixRhsFinalize(rhs))


def rhs_array(rhs):
    """A copy of the vectors a descriptor holds (permuted numbering)."""
    r = ctypes.cast(rhs, ctypes.POINTER(pastix_rhs_s)).contents
    dt = {ENUMS["SpmDouble"]: np.float64, ENUMS["SpmFloat"]: np.float32,
          ENUMS["SpmComplex64"]: np.complex128, ENUMS["SpmComplex32"]: np.complex64}[r.flttype]
    es = np.dtype(dt).itemsize
    buf = (ctypes.c_char * (r.ld * r.n * es)).from_address(r.b)
    a = np.frombuffer(buf, dtype=dt).reshape(r.n, r.ld).T[:r.m]
    return np.array(a, order="F", copy=True)


def subtask_applyorder(pastix_data, direction, b, Bp):
    m = b.shape[0]
    n = 1 if b.ndim == 1 else b.shape[1]
    if b.ndim == 2 and not b.flags.f_contiguous:
        raise ValueError("subtask_applyorder: b must be Fortran-ordered")
    _check("pastix_subtask_applyorder", lib.pastix_subtask_applyorder(pastix_data, direction, m, n, _ptr(b), m, Bp))


def subtask_trsm(pastix_data, side, uplo, trans, diag, b):
    _check("pastix_subtask_trsm", lib.pastix_subtask_trsm(pastix_data, side, uplo, trans, diag, b))


def subtask_diag(pastix_data, b):
    _check("pastix_subtask_diag", lib.pastix_subtask_diag(pastix_data, b))


def subtask_solve(pastix_data, b):
    _check("pastix_subtask_solve", lib.pastix_subtask_solve(pastix_data, b))


def subtask_solve_adv(pastix_data, transA, b):
    _check("pastix_subtask_solve_adv", lib.pastix_subtask_solve_adv(pastix_data, transA, b))


def subtask_refine(pastix_data, b, x):
    _check("pastix_subtask_refine", lib.pastix_subtask_refine(pastix_data, b, x))


def setSchurUnknownList(pastix_data, schur_list):
    schur_list = np.ascontiguousarray(schur_list, dtype=np.int64)
    lib.pastixSetSchurUnknownList(pastix_data, schur_list.shape[0], _ptr(schur_list))


def getSchur(pastix_data, S):
    """pastixGetSchur into the Fortran-ordered square array S."""
    if not S.flags.f_contiguous:
        raise ValueError("getSchur: S must be Fortran-ordered")
    _check("pastixGetSchur", lib.pastixGetSchur(pastix_data, _ptr(S), S.shape[0]))


def rhsSchurGet(pastix_data, rhsB, B):
    m = B.shape[0]
    n = 1 if B.ndim == 1 else B.shape[1]
    _check("pastixRhsSchurGet", lib.pastixRhsSchurGet(pastix_data, m, n, rhsB, _ptr(B), m))


def rhsSchurSet(pastix_data, B, rhsB):
    m = B.shape[0]
    n = 1 if B.ndim == 1 else B.shape[1]
    B = np.asfortranarray(B)
    _check("pastixRhsSchurSet", lib.pastixRhsSchurSet(pastix_data, m, n, _ptr(B), m, rhsB))


def getDiag(pastix_data, n, dtype=np.float64, incx=1):
    """pastixGetDiag: the factor's diagonal in the permuted numbering."""
    out = np.zeros(n * incx, dtype=dtype)
    _check("pastixGetDiag", lib.pastixGetDiag(pastix_data, _ptr(out), incx))
    return out[::incx]


def set_option(pastix_data, key, value):
    _check("pastix_amd_set_option", lib.pastix_amd_set_option(pastix_data, key.encode(), float(value)))


def device_count():
    return lib.pxd_device_count()


# ---------------------------------------------------------------------------
# Engine introspection (structure and factor export for the parity tests)

def solver_export(pastix_data):
    c, b, f, n = pastix_int(), pastix_int(), pastix_int(), pastix_int()
    _check("pastix_amd_solver_sizes", lib.pastix_amd_solver_sizes(
        pastix_data, ctypes.byref(c), ctypes.byref(b), ctypes.byref(f), ctypes.byref(n)))
    cb = np.zeros(6 * (c.value + 1), dtype=np.int64)
    bk = np.zeros(5 * b.value, dtype=np.int64)
    bt = np.zeros(max(1, b.value - c.value), dtype=np.int64)
    _check("pastix_amd_solver_export", lib.pastix_amd_solver_export(pastix_data, _ptr(cb), _ptr(bk), _ptr(bt)))
    return dict(cblknbr=c.value, bloknbr=b.value, coefnbr=f.value, nodenbr=n.value,
                cblk=cb, blok=bk, browtab=bt)


def order_export(pastix_data):
    v, c = pastix_int(), pastix_int()
    _check("pastix_amd_order_sizes", lib.pastix_amd_order_sizes(pastix_data, ctypes.byref(v), ctypes.byref(c)))
    perm = np.zeros(v.value, dtype=np.int64)
    peri = np.zeros(v.value, dtype=np.int64)
    rang = np.zeros(c.value + 1, dtype=np.int64)
    tree = np.zeros(c.value, dtype=np.int64)
    _check("pastix_amd_order_export", lib.pastix_amd_order_export(pastix_data, _ptr(perm), _ptr(peri),
                                                                   _ptr(rang), _ptr(tree)))
    return perm, peri, rang, tree


def symbol_export(pastix_data):
    c, b = pastix_int(), pastix_int()
    _check("pastix_amd_symbol_sizes", lib.pastix_amd_symbol_sizes(pastix_data, ctypes.byref(c), ctypes.byref(b)))
    cb = np.zeros(3 * (c.value + 1), dtype=np.int64)
    bk = np.zeros(4 * b.value, dtype=np.int64)
    _check("pastix_amd_symbol_export", lib.pastix_amd_symbol_export(pastix_data, _ptr(cb), _ptr(bk)))
    return cb.reshape(-1, 3), bk.reshape(-1, 4)


def bcsc_export(pastix_data, complex_=False):
    n, nnz = pastix_int(), pastix_int()
    _check("pastix_amd_bcsc_sizes", lib.pastix_amd_bcsc_sizes(pastix_data, ctypes.byref(n), ctypes.byref(nnz)))
    cp = np.zeros(n.value + 1, dtype=np.int64)
    rt = np.zeros(nnz.value, dtype=np.int64)
    v = np.zeros(nnz.value, dtype=np.complex128 if complex_ else np.float64)
    _check("pastix_amd_bcsc_export", lib.pastix_amd_bcsc_export(pastix_data, _ptr(cp), _ptr(rt), _ptr(v)))
    return cp, rt, v


def get_coeftab(pastix_data, coefnbr, dtype, side=0):
    out = np.zeros(coefnbr, dtype=dtype)
    _check("pastix_amd_get_coeftab", lib.pastix_amd_get_coeftab(pastix_data, side, _ptr(out), out.nbytes))
    return out


STAT_NAMES = ["fact_seconds", "fill_seconds", "panel_seconds", "trsm_seconds", "update_seconds",
              "solve_seconds", "levels", "update_tiles", "launches", "coef_bytes"]


def get_stats(pastix_data):
    v = np.zeros(len(STAT_NAMES))
    _check("pastix_amd_get_stats", lib.pastix_amd_get_stats(pastix_data, _ptr(v), len(v)))
    return dict(zip(STAT_NAMES, v.tolist()))
