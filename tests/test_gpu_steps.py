"""The solve subtasks and the SpMV on the GPU against the oracle.

- pastix_subtask_trsm / pastix_subtask_diag, one sweep at a time, against
  the oracle's sweeps (orc_solve_step, kernels/solve_ztrsmsp.c) on the
  device's own factor, and their composition against pastix_subtask_solve
  (reference sopalin/pastix_task_solve.c:510-677 calls exactly these steps);
- the transposed LU solve (op(A) = A^T: U^T y = b, L^T x = y);
- the full device solve in single precision and with LU factors against the
  oracle's solve on the same factor (the double Cholesky / LDL^T case is in
  test_gpu_solve.py);
- pastixGetDiag against the factor's diagonal tiles;
- the device SpMV against the oracle's, bit for bit on integer-valued data
  (symmetric, general through the stored transpose, complex).
"""
import ctypes

import numpy as np
import pytest

import pastix_amd as px
import oracle_ref as orc
from helpers import analyzed, make_problem, permute_rhs

pytestmark = pytest.mark.gpu

CASES = [
    # kind, grid edge, factorization, mixed
    ("lap7", 9, px.PastixFactPOTRF, 0),
    ("lap7", 9, px.PastixFactPOTRF, 1),
    ("lap7sd", 8, px.PastixFactLDLT, 0),
    ("lap7sd", 8, px.PastixFactLDLT, 1),
    ("convdiff27", 7, px.PastixFactLU, 0),
    ("convdiff27", 7, px.PastixFactLU, 1),
    ("helmholtz7", 7, px.PastixFactLDLT, 0),
    ("helmholtz7", 7, px.PastixFactLDLT, 1),
    ("helmholtz7", 6, px.PastixFactLU, 0),
    ("helmholtz7", 6, px.PastixFactLU, 1),
]


def _factorized(kind, nx, facto, mixed):
    spm = make_problem(kind, nx)
    pd, iparm, dparm = analyzed(spm, nx, facto=facto, mixed=mixed)
    px.task_numfact(pd, spm)
    s = px.solver_export(pd)
    cplx = spm.is_complex
    fdt = {(False, 0): np.float64, (False, 1): np.float32, (True, 0): np.complex128, (True, 1): np.complex64}
    dt = fdt[(cplx, mixed)]
    L = px.get_coeftab(pd, s["coefnbr"], dt)
    U = px.get_coeftab(pd, s["coefnbr"], dt, side=1) if facto == px.PastixFactLU else None
    return spm, pd, s, L, U


def _rhs(spm, nrhs, seed):
    g = np.random.default_rng(seed)
    b = g.uniform(-1, 1, (spm.n, nrhs))
    if spm.is_complex:
        b = b + 1j * g.uniform(-1, 1, (spm.n, nrhs))
    return np.asfortranarray(b)


def _apply(pd, bp, fn):
    """Run fn on a right-hand-side descriptor holding bp (permuted numbering)."""
    Bp = px.rhsInit()
    # applyorder Forward with the identity: scatter bp into the descriptor
    # through the inverse permutation, so the descriptor holds bp itself
    perm = px.order_export(pd)[0]
    b = np.asfortranarray(bp[perm])
    px.subtask_applyorder(pd, px.PastixDirForward, b, Bp)
    fn(Bp)
    out = px.rhs_array(Bp)
    px.rhsFinalize(Bp)
    return out


def _tol(dtype, x):
    return 300 * np.finfo(dtype).eps * max(1.0, np.abs(x).max())


@pytest.mark.parametrize("kind,nx,facto,mixed", CASES)
def test_trsm_steps_match_oracle_and_compose_into_the_solve(kind, nx, facto, mixed):
    spm, pd, s, L, U = _factorized(kind, nx, facto, mixed)
    eps_t = np.float32 if mixed else np.float64
    bp = _rhs(spm, 3, 1)
    lu = facto == px.PastixFactLU
    unit = facto != px.PastixFactPOTRF
    diag_l = px.PastixUnit if unit else px.PastixNonUnit
    L_ = px.PastixLeft
    # forward L y = b
    y = _apply(pd, bp, lambda B: px.subtask_trsm(pd, L_, px.PastixLower, px.PastixNoTrans, diag_l, B))
    yo = orc.solve_step(s, L, orc.STEP_FWD, unit, bp)
    assert np.abs(y - yo).max() <= _tol(eps_t, yo)
    if facto == px.PastixFactLDLT:
        z = _apply(pd, y, lambda B: px.subtask_diag(pd, B))
        zo = orc.solve_step(s, L, orc.STEP_DIAG, False, y)
        assert np.abs(z - zo).max() <= _tol(eps_t, zo)
    else:
        z = y
    if lu:
        x = _apply(pd, z, lambda B: px.subtask_trsm(pd, L_, px.PastixUpper, px.PastixNoTrans, px.PastixNonUnit, B))
        xo = orc.solve_step(s, U, orc.STEP_BWD, False, z)
    else:
        x = _apply(pd, z, lambda B: px.subtask_trsm(pd, L_, px.PastixLower, px.PastixTrans, diag_l, B))
        xo = orc.solve_step(s, L, orc.STEP_BWD, unit, z)
    assert np.abs(x - xo).max() <= _tol(eps_t, xo)
    # the steps in a row reproduce the full subtask solve
    xs = _apply(pd, bp, lambda B: px.subtask_solve(pd, B))
    assert np.abs(x - xs).max() <= 10 * np.finfo(eps_t).eps * max(1.0, np.abs(xs).max())
    # the full device solve against the oracle's on the device factor
    # (single precision and LU included)
    assert np.abs(xs - xo).max() <= _tol(eps_t, xo)
    px.finalize(pd)


@pytest.mark.parametrize("mixed", [0, 1])
@pytest.mark.parametrize("kind,nx", [("convdiff27", 7), ("helmholtz7", 6)])
def test_transposed_lu_solve(kind, nx, mixed):
    spm, pd, s, L, U = _factorized(kind, nx, px.PastixFactLU, mixed)
    eps_t = np.float32 if mixed else np.float64
    bp = _rhs(spm, 2, 4)
    x = _apply(pd, bp, lambda B: px.subtask_solve_adv(pd, px.PastixTrans, B))
    xo = orc.solve_step(s, L, orc.STEP_BWD, True, orc.solve_step(s, U, orc.STEP_FWD, False, bp))
    assert np.abs(x - xo).max() <= _tol(eps_t, xo)
    # the same through the single steps U^T then L^T
    y = _apply(pd, bp, lambda B: px.subtask_trsm(pd, px.PastixLeft, px.PastixUpper, px.PastixTrans,
                                                 px.PastixNonUnit, B))
    x2 = _apply(pd, y, lambda B: px.subtask_trsm(pd, px.PastixLeft, px.PastixLower, px.PastixTrans,
                                                 px.PastixUnit, B))
    assert np.abs(x2 - x).max() <= 10 * np.finfo(eps_t).eps * max(1.0, np.abs(x).max())
    if not mixed:
        # A^T x = b in the original numbering
        cp, rt, v = px.bcsc_export(pd, complex_=spm.is_complex)
        import scipy.sparse as sp
        Ap = sp.csc_matrix((v, rt, cp), shape=(spm.n, spm.n))
        r = Ap.T @ x - bp
        assert np.abs(r).max() <= 1e-12 * np.abs(bp).max() * 1e2
    px.finalize(pd)


def test_trsm_rejects_sweeps_the_factorization_does_not_have():
    spm, pd, s, L, U = _factorized("lap7", 6, px.PastixFactPOTRF, 0)
    bp = _rhs(spm, 1, 0)
    for uplo, trans, diag in ((px.PastixUpper, px.PastixNoTrans, px.PastixNonUnit),
                              (px.PastixLower, px.PastixNoTrans, px.PastixUnit)):
        with pytest.raises(px.PastixError) as e:
            _apply(pd, bp, lambda B: px.subtask_trsm(pd, px.PastixLeft, uplo, trans, diag, B))
        assert e.value.code == px.PASTIX_ERR_BADPARAMETER
    with pytest.raises(px.PastixError):
        _apply(pd, bp, lambda B: px.subtask_trsm(pd, px.PastixRight, px.PastixLower, px.PastixNoTrans,
                                                 px.PastixNonUnit, B))
    px.finalize(pd)


@pytest.mark.parametrize("kind,nx,facto,mixed", [CASES[0], CASES[1], CASES[2], CASES[4], CASES[7]])
def test_get_diag_returns_the_factor_diagonal(kind, nx, facto, mixed):
    spm, pd, s, L, U = _factorized(kind, nx, facto, mixed)
    cb = s["cblk"].reshape(-1, 6)
    want = np.zeros(s["nodenbr"], dtype=L.dtype)
    for k in range(s["cblknbr"]):
        nk = cb[k, 1] - cb[k, 0] + 1
        want[cb[k, 0]:cb[k, 0] + nk] = L[cb[k, 4]:cb[k, 4] + nk * nk:nk + 1]
    dt = spm.dtype
    got = px.getDiag(pd, s["nodenbr"], dt, incx=2)
    assert np.array_equal(got, want.astype(dt))
    px.finalize(pd)


def _integer_spm(kind, nx, seed, cplx=False):
    base = make_problem(kind, nx)
    cp, rt, v = base.arrays()
    g = np.random.default_rng(seed)
    vi = g.integers(-9, 10, size=v.shape).astype(np.float64)
    if cplx:
        vi = vi + 1j * g.integers(-9, 10, size=v.shape)
    # keep the diagonal dominant (irrelevant to the SpMV, harmless to the analysis)
    return px.Spm.from_csc(base.n, cp, rt, vi, mtxtype=base.s.mtxtype)


@pytest.mark.parametrize("nrhs", [1, 3, 9])
@pytest.mark.parametrize("kind,cplx", [("lap7", False), ("convdiff27", False), ("lap7", True)])
def test_device_spmv_is_exact_on_integer_data(kind, cplx, nrhs):
    """y = alpha A x + beta y with integers everywhere: every partial sum is
    exact in double, so the device's gather order cannot matter and the
    result equals the oracle's SpMV (bcsc/bcsc_zspmv.c:629-678) bit for bit."""
    nx = 8
    spm = _integer_spm(kind, nx, 5, cplx)
    pd, iparm, dparm = analyzed(spm, nx)
    px._check("spm2bcsc", px.lib.pastix_subtask_spm2bcsc(pd, ctypes.byref(spm.s)))
    px._check("bcsc2ctab", px.lib.pastix_subtask_bcsc2ctab(pd))
    n = spm.n
    g = np.random.default_rng(7)
    dt = spm.dtype
    x = g.integers(-30, 31, size=(n, nrhs)).astype(dt)
    y = g.integers(-30, 31, size=(n, nrhs)).astype(dt)
    if cplx:
        x = x + 1j * g.integers(-30, 31, size=(n, nrhs))
        y = y + 1j * g.integers(-30, 31, size=(n, nrhs))
    x = np.asfortranarray(x)
    y = np.asfortranarray(y)
    dev = ctypes.c_void_p()
    px._check("device", px.lib.pastix_amd_device_handle(pd, ctypes.byref(dev)))
    dx, dy = ctypes.c_void_p(), ctypes.c_void_p()
    px._check("malloc", px.lib.pxd_malloc(dev, x.nbytes, ctypes.byref(dx)))
    px._check("malloc", px.lib.pxd_malloc(dev, y.nbytes, ctypes.byref(dy)))
    px._check("h2d", px.lib.pxd_h2d(dev, dx, x.ctypes.data_as(ctypes.c_void_p), x.nbytes))
    px._check("h2d", px.lib.pxd_h2d(dev, dy, y.ctypes.data_as(ctypes.c_void_p), y.nbytes))
    px._check("pxd_spmv", px.lib.pxd_spmv(dev, 2.0, dx, -3.0, dy, nrhs, n))
    out = np.zeros_like(y)
    px._check("d2h", px.lib.pxd_d2h(dev, out.ctypes.data_as(ctypes.c_void_p), dy, y.nbytes))
    px.lib.pxd_free(dev, dx)
    px.lib.pxd_free(dev, dy)
    # the oracle on the same permuted full-pattern CSC (for a general matrix,
    # rows of A are the columns of its transpose)
    cp, rt, v = px.bcsc_export(pd, complex_=cplx)
    if kind == "convdiff27":
        cp, rt, v = orc.transpose_csc(cp, rt, v)
    want = np.zeros_like(y)
    for q in range(nrhs):
        want[:, q] = orc.spmv(cp, rt, v, x[:, q], 2.0, -3.0, y[:, q])
    assert np.array_equal(out, want)
    px.finalize(pd)
