"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure: the oracle is the checker for the device engine and the
CPU baseline of the benchmark; the product never loads it.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

_lib = None
_p = ctypes.c_void_p
_i64 = ctypes.c_int64


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_PATH):
            raise ImportError("oracle library missing: run `python -m pastix_amd.build`")
        _lib = ctypes.CDLL(ORACLE_PATH)
        _lib.orc_factorize.restype = ctypes.c_int
        _lib.orc_factorize.argtypes = [ctypes.c_int, _i64, _i64, _p, _p, _i64, _p, _p, _p,
                                       ctypes.c_double, _p, ctypes.c_int, _p, _p]
        _lib.orc_solve.restype = ctypes.c_int
        _lib.orc_solve.argtypes = [ctypes.c_int, _i64, _i64, _p, _p, _p, _p, _i64, _p, _i64]
        _lib.orc_spmv.restype = ctypes.c_int
        _lib.orc_spmv.argtypes = [_i64, _p, _p, _p, ctypes.c_double, _p, ctypes.c_double, _p]
        _lib.orc_gmres.restype = _i64
        _lib.orc_gmres.argtypes = [ctypes.c_int, _i64, _i64, _p, _p, _p, _p, _p, _p, _p, _p, _p,
                                   _i64, _i64, ctypes.c_double]
        _lib.orc_blas.restype = ctypes.c_char_p
        _lib.orc_factorize_lu.restype = ctypes.c_int
        _lib.orc_factorize_lu.argtypes = [ctypes.c_int, _i64, _i64, _p, _p, _i64, _p, _p, _p, _p, _p, _p,
                                          ctypes.c_double, _p, _p, _p]
        _lib.orc_factorize_ldlt.restype = ctypes.c_int
        _lib.orc_factorize_ldlt.argtypes = [ctypes.c_int, _i64, _i64, _p, _p, _i64, _p, _p, _p,
                                            ctypes.c_double, _p, _p]
        _lib.orc_solve_ldlt.restype = ctypes.c_int
        _lib.orc_solve_ldlt.argtypes = [ctypes.c_int, _i64, _i64, _p, _p, _p, _p, _i64, _p, _i64]
        _lib.orc_solve_step.restype = ctypes.c_int
        _lib.orc_solve_step.argtypes = [ctypes.c_int, _i64, _i64, _p, _p, _p, _p, ctypes.c_int, ctypes.c_int,
                                        _i64, _p, _i64]
        _lib.orc_spmv_z.restype = ctypes.c_int
        _lib.orc_spmv_z.argtypes = [_i64, _p, _p, _p, ctypes.c_double, _p, ctypes.c_double, _p]
        _lib.orc_solve_lu.restype = ctypes.c_int
        _lib.orc_solve_lu.argtypes = [ctypes.c_int, _i64, _i64, _p, _p, _p, _p, _p, _i64, _p, _i64]
    return _lib


def _ptr(a):
    return a.ctypes.data_as(_p)


def blas_name():
    return lib().orc_blas().decode()


class OracleFactor:
    """Cholesky factor computed by the oracle on an exported solver structure."""

    def __init__(self, solver, bcsc, thresh, dtype=np.float64, nthreads=1):
        self.s = solver
        self.cp, self.rt, self.val = bcsc
        self.dtype = dtype
        self.flt = 3 if dtype == np.float64 else 2
        self.n = solver["nodenbr"]
        self.L = np.zeros(solver["coefnbr"], dtype=dtype)
        npiv = np.zeros(1, dtype=np.int64)
        sec = np.zeros(1)
        lib().orc_factorize(self.flt, self.n, solver["cblknbr"], _ptr(solver["cblk"]), _ptr(solver["blok"]),
                            solver["coefnbr"], _ptr(self.cp), _ptr(self.rt), _ptr(self.val), thresh,
                            _ptr(self.L), nthreads, _ptr(npiv), _ptr(sec))
        self.npiv = int(npiv[0])
        self.seconds = float(sec[0])

    def solve(self, b):
        x = np.array(b, dtype=np.float64, order="F", copy=True)
        nrhs = 1 if x.ndim == 1 else x.shape[1]
        lib().orc_solve(self.flt, self.n, self.s["cblknbr"], _ptr(self.s["cblk"]), _ptr(self.s["blok"]),
                        _ptr(self.s["browtab"]), _ptr(self.L), nrhs, _ptr(x), self.n)
        return x

    def gmres(self, b, x0, im=25, itermax=250, eps=1e-12):
        x = np.array(x0, dtype=np.float64, copy=True)
        b = np.ascontiguousarray(b, dtype=np.float64)
        it = lib().orc_gmres(self.flt, self.n, self.s["cblknbr"], _ptr(self.s["cblk"]), _ptr(self.s["blok"]),
                             _ptr(self.s["browtab"]), _ptr(self.L), _ptr(self.cp), _ptr(self.rt),
                             _ptr(self.val), _ptr(b), _ptr(x), im, itermax, eps)
        return x, int(it)


def transpose_csc(cp, rt, val):
    """CSC of A^T from the CSC of A (rows sorted in every column)."""
    import scipy.sparse as sp
    n = len(cp) - 1
    T = sp.csc_matrix((val, rt, cp), shape=(n, n)).T.tocsc()
    T.sort_indices()
    return T.indptr.astype(np.int64), T.indices.astype(np.int64), np.ascontiguousarray(T.data)


_FLT = {np.dtype(np.float64): 3, np.dtype(np.float32): 2, np.dtype(np.complex128): 5, np.dtype(np.complex64): 4}


class OracleLU:
    """LU factor with static pivoting computed by the oracle: L tiles (unit
    lower + the U part of the diagonal tiles) and U^T tiles."""

    def __init__(self, solver, bcsc, thresh, dtype=np.float64):
        self.s = solver
        self.cp, self.rt, self.val = bcsc
        self.tcp, self.trt, self.tval = transpose_csc(*bcsc)
        self.dtype = dtype
        self.flt = _FLT[np.dtype(dtype)]
        self.n = solver["nodenbr"]
        self.L = np.zeros(solver["coefnbr"], dtype=dtype)
        self.U = np.zeros(solver["coefnbr"], dtype=dtype)
        npiv = np.zeros(1, dtype=np.int64)
        lib().orc_factorize_lu(self.flt, self.n, solver["cblknbr"], _ptr(solver["cblk"]), _ptr(solver["blok"]),
                               solver["coefnbr"], _ptr(self.cp), _ptr(self.rt), _ptr(self.val), _ptr(self.tcp),
                               _ptr(self.trt), _ptr(self.tval), thresh, _ptr(self.L), _ptr(self.U), _ptr(npiv))
        self.npiv = int(npiv[0])

    def solve(self, b):
        x = np.array(b, dtype=np.complex128 if self.flt in (4, 5) else np.float64, order="F", copy=True)
        nrhs = 1 if x.ndim == 1 else x.shape[1]
        lib().orc_solve_lu(self.flt, self.n, self.s["cblknbr"], _ptr(self.s["cblk"]), _ptr(self.s["blok"]),
                           _ptr(self.s["browtab"]), _ptr(self.L), _ptr(self.U), nrhs, _ptr(x), self.n)
        return x


class OracleLDLT:
    """LDL^T factor with static pivoting computed by the oracle: the diagonal
    tiles hold L (strictly below), D and D L^T; the other tiles hold L."""

    def __init__(self, solver, bcsc, thresh, dtype=np.float64):
        self.s = solver
        self.cp, self.rt, self.val = bcsc
        self.dtype = dtype
        self.flt = _FLT[np.dtype(dtype)]
        self.n = solver["nodenbr"]
        self.L = np.zeros(solver["coefnbr"], dtype=dtype)
        npiv = np.zeros(1, dtype=np.int64)
        lib().orc_factorize_ldlt(self.flt, self.n, solver["cblknbr"], _ptr(solver["cblk"]), _ptr(solver["blok"]),
                                 solver["coefnbr"], _ptr(self.cp), _ptr(self.rt), _ptr(self.val), thresh,
                                 _ptr(self.L), _ptr(npiv))
        self.npiv = int(npiv[0])

    def solve(self, b):
        x = np.array(b, dtype=np.complex128 if self.flt in (4, 5) else np.float64, order="F", copy=True)
        nrhs = 1 if x.ndim == 1 else x.shape[1]
        lib().orc_solve_ldlt(self.flt, self.n, self.s["cblknbr"], _ptr(self.s["cblk"]), _ptr(self.s["blok"]),
                             _ptr(self.s["browtab"]), _ptr(self.L), nrhs, _ptr(x), self.n)
        return x


def spmv(cp, rt, val, x, alpha=1.0, beta=0.0, y=None):
    """y = alpha A x + beta y on a full-pattern CSC (column j is row j of the
    operator), real or complex double."""
    n = len(cp) - 1
    cplx = np.iscomplexobj(val)
    dt = np.complex128 if cplx else np.float64
    x = np.ascontiguousarray(x, dtype=dt)
    val = np.ascontiguousarray(val, dtype=dt)
    y = np.zeros(n, dtype=dt) if y is None else np.array(y, dtype=dt, copy=True)
    f = lib().orc_spmv_z if cplx else lib().orc_spmv
    f(n, _ptr(cp), _ptr(rt), _ptr(val), alpha, _ptr(x), beta, _ptr(y))
    return y


# solve steps of orc_solve_step: forward sweep, backward (transposed) sweep,
# division by the diagonal of the diagonal tiles
STEP_FWD, STEP_BWD, STEP_DIAG = 0, 1, 2


def solve_step(solver, A, step, unit, b):
    """One sweep of a solve (kernels/solve_ztrsmsp.c) on the tiles A (any of
    the four arithmetics); b double or complex double, returned solved."""
    flt = _FLT[np.dtype(A.dtype)]
    x = np.array(b, dtype=np.complex128 if flt in (4, 5) else np.float64, order="F", copy=True)
    nrhs = 1 if x.ndim == 1 else x.shape[1]
    lib().orc_solve_step(flt, solver["nodenbr"], solver["cblknbr"], _ptr(solver["cblk"]), _ptr(solver["blok"]),
                         _ptr(solver["browtab"]), _ptr(np.ascontiguousarray(A)), step, int(unit), nrhs, _ptr(x),
                         solver["nodenbr"])
    return x


def tiles(solver, L):
    """Yield (cblk, blok, tile ndarray view) for every blok of the factor."""
    cb = solver["cblk"].reshape(-1, 6)
    bk = solver["blok"].reshape(-1, 5)
    for k in range(solver["cblknbr"]):
        n = cb[k, 1] - cb[k, 0] + 1
        for b in range(cb[k, 2], cb[k + 1, 2]):
            rows = bk[b, 1] - bk[b, 0] + 1
            off = cb[k, 4] + bk[b, 4]
            yield k, b, L[off:off + rows * n].reshape(n, rows).T


def dense_from_factor(solver, L):
    """Assemble the dense lower factor (permuted numbering) from the tiles."""
    n = solver["nodenbr"]
    D = np.zeros((n, n), dtype=L.dtype)
    cb = solver["cblk"].reshape(-1, 6)
    bk = solver["blok"].reshape(-1, 5)
    for k, b, t in tiles(solver, L):
        fc = cb[k, 0]
        fr = bk[b, 0]
        D[fr:fr + t.shape[0], fc:fc + t.shape[1]] = t
    return np.tril(D)


# ---------------------------------------------------------------------------
# Learned cost model: float64 restatement of the perceptron whose forward and
# backward passes run on the GPU (10 inputs, two tanh layers of 32 units, one
# linear output; parameters W1[10][32], b1, W2[32][32], b2, w3, b3 in one array).

MLP_F, MLP_H = 10, 32


def mlp_unpack(p):
    p = np.asarray(p, dtype=np.float64)
    F, H = MLP_F, MLP_H
    o = 0
    W1 = p[o:o + F * H].reshape(F, H); o += F * H
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o]; o += 1
    assert o == p.size
    return W1, b1, W2, b2, w3, b3


def mlp_forward(p, X):
    W1, b1, W2, b2, w3, b3 = mlp_unpack(p)
    h1 = np.tanh(X @ W1 + b1)
    h2 = np.tanh(h1 @ W2 + b2)
    return h2 @ w3 + b3


def mlp_loss_grad(p, X, t):
    """Mean squared error sum_r (y_r - t_r)^2 / n and its gradient."""
    X = np.asarray(X, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    n = X.shape[0]
    W1, b1, W2, b2, w3, b3 = mlp_unpack(p)
    h1 = np.tanh(X @ W1 + b1)
    h2 = np.tanh(h1 @ W2 + b2)
    y = h2 @ w3 + b3
    e = y - t
    dy = 2.0 * e / n
    g2 = np.outer(dy, w3) * (1.0 - h2 ** 2)
    g1 = (g2 @ W2.T) * (1.0 - h1 ** 2)
    grad = np.concatenate([(X.T @ g1).ravel(), g1.sum(0), (h1.T @ g2).ravel(), g2.sum(0),
                           h2.T @ dy, [dy.sum()]])
    return float((e ** 2).sum() / n), grad
