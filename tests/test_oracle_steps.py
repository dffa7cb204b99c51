"""The oracle's solve steps, SpMV and the single-precision LU / LDL^T
factorizations pinned against dense linear algebra, and the spm helpers
(spmBackwardError, spmMatMat, spmCheckAxb) against numpy restatements of
their formulas.

The solve steps (orc_solve_step, a restatement of kernels/solve_ztrsmsp.c
forward / backward / diag) are what pastix_subtask_trsm and
pastix_subtask_diag apply one at a time; here each is compared with a dense
triangular solve of the matrix the tiles assemble into, and their
compositions with the oracle's full solves."""
import ctypes

import numpy as np
import pytest
import scipy.linalg as la
import scipy.sparse as sp

import pastix_amd as px
import oracle_ref as orc
from helpers import analyzed, make_problem, static_pivot_threshold


def _dense(cp, rt, v, n):
    A = np.zeros((n, n), dtype=v.dtype)
    for j in range(n):
        A[rt[cp[j]:cp[j + 1]], j] = v[cp[j]:cp[j + 1]]
    return A


def _bcsc(kind, dims, facto, complex_=False):
    spm = make_problem(kind, *dims)
    pd, iparm, dparm = analyzed(spm, *dims, facto=facto)
    px._check("spm2bcsc", px.lib.pastix_subtask_spm2bcsc(pd, ctypes.byref(spm.s)))
    return spm, pd, dparm, px.solver_export(pd), px.bcsc_export(pd, complex_=complex_)


def _rand(n, nrhs, cplx, seed):
    g = np.random.default_rng(seed)
    b = g.uniform(-1, 1, (n, nrhs))
    if cplx:
        b = b + 1j * g.uniform(-1, 1, (n, nrhs))
    return np.asfortranarray(b)


CASES = [
    # kind, dims, factorization, factor dtype
    ("lap7", (5, 4, 4), "PastixFactPOTRF", np.float64),
    ("lap7", (5, 4, 4), "PastixFactPOTRF", np.float32),
    ("convdiff27", (4, 4, 3), "PastixFactLU", np.float64),
    ("convdiff27", (4, 4, 3), "PastixFactLU", np.float32),
    ("helmholtz7", (4, 4, 4), "PastixFactLU", np.complex128),
    ("helmholtz7", (4, 4, 4), "PastixFactLDLT", np.complex128),
    ("helmholtz7", (4, 4, 4), "PastixFactLDLT", np.complex64),
    ("lap7sd", (5, 4, 3), "PastixFactLDLT", np.float64),
    ("lap7sd", (5, 4, 3), "PastixFactLDLT", np.float32),
]


def _factor(kind, dims, facto, dtype):
    cplx = np.iscomplexobj(np.zeros(1, dtype))
    spm, pd, dparm, s, bc = _bcsc(kind, dims, getattr(px, facto), complex_=cplx)
    tau = static_pivot_threshold(dparm, np.float32 if dtype in (np.float32, np.complex64) else np.float64)
    if facto == "PastixFactPOTRF":
        F = orc.OracleFactor(s, bc, tau, dtype)
    elif facto == "PastixFactLU":
        F = orc.OracleLU(s, bc, tau, dtype)
    else:
        F = orc.OracleLDLT(s, bc, tau, dtype)
    return spm, pd, s, bc, F


@pytest.mark.parametrize("kind,dims,facto,dtype", CASES)
def test_solve_steps_equal_dense_triangular_solves(kind, dims, facto, dtype):
    spm, pd, s, bc, F = _factor(kind, dims, facto, dtype)
    n = s["nodenbr"]
    cplx = np.iscomplexobj(F.L)
    wide = np.complex128 if cplx else np.float64
    b = _rand(n, 3, cplx, 11)
    Ld = orc.dense_from_factor(s, F.L).astype(wide)
    unit = facto != "PastixFactPOTRF"
    Lm = np.tril(Ld, -1) + np.eye(n) if unit else Ld
    tol = 1e3 * np.finfo(dtype).eps * np.linalg.cond(Lm)
    # forward: L y = b
    y = orc.solve_step(s, F.L, orc.STEP_FWD, unit, b)
    assert np.abs(y - la.solve_triangular(Lm, b, lower=True)).max() <= tol * np.abs(y).max()
    # backward: L^T x = b
    x = orc.solve_step(s, F.L, orc.STEP_BWD, unit, b)
    assert np.abs(x - la.solve_triangular(Lm.T, b, lower=False)).max() <= tol * np.abs(x).max()
    # diagonal: b / diag(L tiles)
    d = orc.solve_step(s, F.L, orc.STEP_DIAG, False, b)
    assert np.abs(d - b / np.diag(Ld)[:, None]).max() <= 4 * np.finfo(dtype).eps * np.abs(d).max()
    if facto == "PastixFactLU":
        # the U^T tiles: non-unit lower triangles (U_kk^T) and U^T off the diagonal
        Um = orc.dense_from_factor(s, F.U).astype(wide)
        yu = orc.solve_step(s, F.U, orc.STEP_FWD, False, b)
        assert np.abs(yu - la.solve_triangular(Um, b, lower=True)).max() <= tol * np.abs(yu).max() * 10
        xu = orc.solve_step(s, F.U, orc.STEP_BWD, False, b)
        assert np.abs(xu - la.solve_triangular(Um.T, b, lower=False)).max() <= tol * np.abs(xu).max() * 10
    px.finalize(pd)


@pytest.mark.parametrize("kind,dims,facto,dtype", CASES)
def test_solve_steps_compose_into_the_full_solve(kind, dims, facto, dtype):
    """forward, [diag], backward reproduce the oracle's own full solve, and the
    swapped LU sides solve with A^T."""
    spm, pd, s, bc, F = _factor(kind, dims, facto, dtype)
    n = s["nodenbr"]
    cplx = np.iscomplexobj(F.L)
    b = _rand(n, 2, cplx, 5)
    unit = facto != "PastixFactPOTRF"
    if facto == "PastixFactLU":
        y = orc.solve_step(s, F.L, orc.STEP_FWD, True, b)
        x = orc.solve_step(s, F.U, orc.STEP_BWD, False, y)
    else:
        y = orc.solve_step(s, F.L, orc.STEP_FWD, unit, b)
        if facto == "PastixFactLDLT":
            y = orc.solve_step(s, F.L, orc.STEP_DIAG, False, y)
        x = orc.solve_step(s, F.L, orc.STEP_BWD, unit, y)
    ref = F.solve(b)
    eps = np.finfo(dtype).eps
    assert np.abs(x - ref).max() <= 100 * eps * np.abs(ref).max()
    if facto == "PastixFactLU" and dtype == np.float64:
        # A^T x = b: U^T y = b, then L^T x = y
        A = _dense(*bc, n)
        yt = orc.solve_step(s, F.U, orc.STEP_FWD, False, b)
        xt = orc.solve_step(s, F.L, orc.STEP_BWD, True, yt)
        assert np.abs(A.T @ xt - b).max() <= 1e-12 * np.abs(b).max() * np.linalg.cond(A)
    px.finalize(pd)


@pytest.mark.parametrize("dims", [(4, 4, 4), (6, 5, 4)])
def test_oracle_single_lu_equals_dense(dims):
    spm, pd, dparm, s, bc = _bcsc("convdiff27", dims, px.PastixFactLU)
    n = s["nodenbr"]
    A = _dense(*bc, n)
    F = orc.OracleLU(s, bc, static_pivot_threshold(dparm, np.float32), np.float32)
    assert F.npiv == 0
    # dense no-pivot LU in double
    D = np.array(A, copy=True)
    for k in range(n):
        D[k + 1:, k] /= D[k, k]
        D[k + 1:, k + 1:] -= np.outer(D[k + 1:, k], D[k, k + 1:])
    packed = np.tril(orc.dense_from_factor(s, F.L).astype(np.float64), -1) + \
        np.triu(orc.dense_from_factor(s, F.U).astype(np.float64).T)
    assert np.abs(packed - D).max() <= 200 * np.finfo(np.float32).eps * np.abs(D).max()
    px.finalize(pd)


@pytest.mark.parametrize("dims", [(4, 4, 4), (6, 5, 4)])
def test_oracle_real_single_ldlt_equals_dense(dims):
    spm, pd, dparm, s, bc = _bcsc("lap7sd", dims, px.PastixFactLDLT)
    n = s["nodenbr"]
    A = _dense(*bc, n)
    F = orc.OracleLDLT(s, bc, static_pivot_threshold(dparm, np.float32), np.float32)
    assert F.npiv == 0
    Lm = np.eye(n)
    d = np.zeros(n)
    W = np.array(A, copy=True)
    for k in range(n):
        d[k] = W[k, k]
        Lm[k + 1:, k] = W[k + 1:, k] / d[k]
        W[k + 1:, k + 1:] -= d[k] * np.outer(Lm[k + 1:, k], Lm[k + 1:, k])
    Lo = orc.dense_from_factor(s, F.L).astype(np.float64)
    assert np.abs(np.tril(Lo, -1) - np.tril(Lm, -1)).max() <= 100 * np.finfo(np.float32).eps
    assert np.abs(np.diag(Lo) - d).max() <= 100 * np.finfo(np.float32).eps * np.abs(d).max()
    px.finalize(pd)


def test_oracle_spmv_exact_on_integers_general_and_complex():
    """Integer-valued entries and vectors: the sums are exact, so the oracle's
    SpMV equals scipy's bit for bit (criterion of test/z_bcsc_tests.c:179-181
    at its tightest)."""
    # general: the 27-point pattern with integer unsymmetric values, through
    # the transposed CSC (row j of A is column j of A^T)
    spm = make_problem("convdiff27", 4, 4, 3)
    cp, rt, v = spm.arrays()
    n = spm.n
    g = np.random.default_rng(3)
    vi = g.integers(-9, 10, size=v.shape).astype(np.float64)
    M = sp.csc_matrix((vi, rt, cp), shape=(n, n))
    T = M.T.tocsc()
    T.sort_indices()
    x = g.integers(-50, 51, size=n).astype(np.float64)
    y = orc.spmv(T.indptr.astype(np.int64), T.indices.astype(np.int64), T.data, x)
    assert np.array_equal(y, M @ x)
    # complex integer values, complex symmetric full pattern
    base = px.Spm.laplacian7(4, 3, 3)
    cp, rt, v = base.arrays()
    n = base.n
    vc = g.integers(-5, 6, size=v.shape) + 1j * g.integers(-5, 6, size=v.shape)
    Mc = sp.csc_matrix((vc, rt, cp), shape=(n, n))
    Mc = ((Mc + Mc.T) / 1).tocsc()
    Mc.sort_indices()
    xc = g.integers(-20, 21, size=n) + 1j * g.integers(-20, 21, size=n)
    yc = orc.spmv(Mc.indptr.astype(np.int64), Mc.indices.astype(np.int64), Mc.data.astype(np.complex128), xc)
    assert np.array_equal(yc, Mc.T @ xc)


def _spm_general(n, density, cplx, seed, mtxtype=None):
    g = np.random.default_rng(seed)
    M = sp.random(n, n, density=density, random_state=seed, format="csc")
    M = M + sp.eye(n) * 4
    if cplx:
        M = M + 1j * sp.random(n, n, density=density, random_state=seed + 1, format="csc")
    M = M.tocsc()
    M.sort_indices()
    vals = M.data.astype(np.complex128 if cplx else np.float64)
    return px.Spm.from_csc(n, M.indptr, M.indices, vals, mtxtype=mtxtype), M.toarray(), g


@pytest.mark.parametrize("cplx", [False, True])
def test_spm_backward_error_and_matmat_match_numpy(cplx):
    """spmBackwardError = max over columns of ||b - A x||_inf / (||A||_inf ||x||_inf
    + ||b||_inf) and spmMatMat = alpha A x + beta y, recomputed with numpy on a
    random general matrix (the formulas the SPM library documents; its code
    is not part of the reference tree)."""
    n = 60
    spm, A, g = _spm_general(n, 0.08, cplx, 7)
    x = np.asfortranarray(g.uniform(-1, 1, (n, 3)) + (1j * g.uniform(-1, 1, (n, 3)) if cplx else 0))
    y0 = np.asfortranarray(g.uniform(-1, 1, (n, 3)) + (1j * g.uniform(-1, 1, (n, 3)) if cplx else 0))
    y = spm.matmat(x, alpha=1.5, beta=-0.5, y=y0)
    assert np.abs(y - (1.5 * A @ x - 0.5 * y0)).max() <= 1e-13 * np.abs(y).max()
    b = np.asfortranarray(A @ x + 1e-9 * y0)
    worst, each = spm.backward_error(b, x)
    nA = np.abs(A).sum(axis=1).max()
    ref = [np.abs(b[:, c] - A @ x[:, c]).max() / (nA * np.abs(x[:, c]).max() + np.abs(b[:, c]).max())
           for c in range(3)]
    assert np.allclose(each, ref, rtol=1e-6, atol=0)
    assert worst == pytest.approx(max(ref), rel=1e-6)
    assert spm.norm("inf") == pytest.approx(nA, rel=1e-14)


def test_spm_symmetric_storage_matmat_matches_numpy():
    """A symmetric spm stores one triangle; spmMatMat applies the full matrix."""
    spm = px.Spm.laplacian7(4, 4, 3)
    cp, rt, v = spm.arrays()
    n = spm.n
    L = sp.csc_matrix((v, rt, cp), shape=(n, n))
    A = (L + L.T - sp.diags(L.diagonal())).toarray()
    x = np.asfortranarray(np.random.default_rng(1).uniform(-1, 1, (n, 2)))
    assert np.abs(spm.matmat(x) - A @ x).max() <= 1e-13


def test_spm_check_axb_reports_the_forward_error():
    """spmCheckAxb with x0 checks ||x - x0||_inf / ||x0||_inf as well as the
    backward error.  On a matrix with one nearly decoupled unknown (diagonal
    1e-6, nothing else in its column), moving that component by 1e-3 leaves a
    backward error near 1e-10 but a forward error of 1e-3: the check passes
    without x0 and fails with it."""
    n = 40
    g = np.random.default_rng(9)
    M = (sp.random(n, n, density=0.1, random_state=9, format="lil") + sp.eye(n) * 4).tolil()
    M[:, 3] = 0
    M[3, 3] = 1e-6
    M = M.tocsc()
    M.sort_indices()
    spm = px.Spm.from_csc(n, M.indptr, M.indices, M.data.astype(np.float64))
    A = M.toarray()
    x0 = np.asfortranarray(g.uniform(-1, 1, (n, 1)))
    b = np.asfortranarray(A @ x0)
    assert px.check_axb(spm, 1e-12, x0, b, np.array(x0, order="F", copy=True)) == 0
    xbad = np.array(x0, order="F", copy=True)
    xbad[3, 0] += 1e-3
    berr, _ = spm.backward_error(b, xbad)
    assert berr < 1e-9
    assert px.check_axb(spm, 1e-9, None, b, xbad) == 0
    assert px.check_axb(spm, 1e-9, x0, b, xbad) != 0
    fe = px.forward_error(x0, xbad, spm)
    assert fe == pytest.approx(1e-3 / np.abs(x0).max(), rel=1e-9)
    assert fe == pytest.approx(px.forward_error(x0, xbad), rel=1e-14)
