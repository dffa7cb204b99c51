"""Host-side checks on the CPU.

- spmMatMat, spmNorm, spmBackwardError, spmForwardError and spmCheckAxb
  against a numpy computation of the same formulas on random real, complex,
  general, symmetric (one triangle stored) and Hermitian CSC matrices: these
  measures judge every solve and refinement test, so they are pinned here
  by dense linear algebra that shares no code with them;
- the Python wrappers refuse right-hand sides whose dtype, shape or layout
  would make the C side read or write memory it does not own;
- a matrix of another arithmetic than the one of the analysis (the analysis
  depends on the pattern only) sets the type of the right-hand sides from
  pastix_subtask_spm2bcsc on, and applyorder sizes the permuted vectors by it
  (reference sopalin/pastix_task_solve.c:197 takes the type from the matrix
  in use).
"""
import ctypes

import numpy as np
import pytest

import pastix_amd as px
from helpers import analyzed


def _random_csc(n, density, kind, cplx, seed):
    """A random CSC matrix (always with its diagonal) and its dense full form.
    kind: "general", "symmetric" or "hermitian" (lower triangle stored)."""
    rng = np.random.default_rng(seed)
    mask = rng.random((n, n)) < density
    np.fill_diagonal(mask, True)
    vals = rng.uniform(-1, 1, (n, n))
    if cplx:
        vals = vals + 1j * rng.uniform(-1, 1, (n, n))
    if kind != "general":
        mask = np.tril(mask)
    if kind == "hermitian":
        np.fill_diagonal(vals, vals.diagonal().real)
    colptr, rows, v = [0], [], []
    for j in range(n):
        idx = np.nonzero(mask[:, j])[0]
        rows.extend(idx.tolist())
        v.extend(vals[idx, j].tolist())
        colptr.append(len(rows))
    S = np.where(mask, vals, 0)
    if kind == "symmetric":
        full = S + np.tril(S, -1).T
    elif kind == "hermitian":
        full = S + np.tril(S, -1).conj().T
    else:
        full = S
    mtx = {"general": px.ENUMS["SpmGeneral"], "symmetric": px.ENUMS["SpmSymmetric"],
           "hermitian": px.ENUMS["SpmHermitian"]}[kind]
    dt = np.complex128 if cplx else np.float64
    spm = px.Spm.from_csc(n, np.array(colptr), np.array(rows), np.array(v, dtype=dt), mtxtype=mtx)
    return spm, full


CASES = [("general", False), ("general", True), ("symmetric", False), ("symmetric", True),
         ("hermitian", True)]


@pytest.mark.parametrize("kind,cplx", CASES)
def test_matmat_and_norms_match_numpy(kind, cplx):
    n = 37
    spm, A = _random_csc(n, 0.15, kind, cplx, seed=3)
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, (n, 3))
    y0 = rng.uniform(-1, 1, (n, 3))
    if cplx:
        x = x + 1j * rng.uniform(-1, 1, (n, 3))
        y0 = y0 + 1j * rng.uniform(-1, 1, (n, 3))
    y = spm.matmat(x, alpha=-0.5, beta=2.0, y=y0)
    np.testing.assert_allclose(y, -0.5 * A @ x + 2.0 * y0, rtol=0, atol=1e-13)
    absA = np.abs(A)
    assert spm.norm("inf") == pytest.approx(absA.sum(axis=1).max(), rel=1e-14)
    assert spm.norm("one") == pytest.approx(absA.sum(axis=0).max(), rel=1e-14)
    assert spm.norm("max") == pytest.approx(absA.max(), rel=1e-15)
    assert spm.norm("fro") == pytest.approx(np.linalg.norm(A, "fro"), rel=1e-13)


@pytest.mark.parametrize("kind,cplx", CASES)
def test_backward_and_forward_errors_match_numpy(kind, cplx):
    n = 41
    spm, A = _random_csc(n, 0.12, kind, cplx, seed=11)
    A = A + n * np.eye(n)   # keep it well conditioned for the solve below
    # rebuild the spm with the shifted diagonal (the diagonal is stored)
    cp, rp, v = spm.arrays()
    for j in range(n):
        for k in range(cp[j], cp[j + 1]):
            if rp[k] == j:
                v[k] += n
    spm = px.Spm.from_csc(n, cp, rp, v, mtxtype=spm.s.mtxtype)
    rng = np.random.default_rng(7)
    x0 = rng.uniform(-1, 1, (n, 4))
    if cplx:
        x0 = x0 + 1j * rng.uniform(-1, 1, (n, 4))
    b = A @ x0
    x = np.linalg.solve(A, b)
    # perturb the columns by different amounts
    pert = rng.uniform(-1, 1, (n, 4)) * np.array([1e-15, 1e-11, 1e-9, 1e-6])
    x = x + pert
    worst, each = spm.backward_error(np.asfortranarray(b), np.asfortranarray(x))
    r = b - A @ x
    nA = np.abs(A).sum(axis=1).max()
    ref = np.abs(r).max(axis=0) / (nA * np.abs(x).max(axis=0) + np.abs(b).max(axis=0))
    # the first column sits at rounding level: its residual depends on the
    # summation order, hence the absolute margin of ~1 ulp of the ratio
    np.testing.assert_allclose(each, ref, rtol=1e-6, atol=1e-16)
    assert worst == pytest.approx(ref.max(), rel=1e-6)
    fe = px.forward_error(x0, x, spm)
    ref_fe = (np.abs(x - x0).max(axis=0) / np.abs(x0).max(axis=0)).max()
    assert fe == pytest.approx(ref_fe, rel=1e-12)
    # spmCheckAxb: backward error within eps; with x0, the forward error too
    eps_ok = 2 * ref.max()
    assert px.check_axb(spm, eps_ok, None, b, x) == 0
    assert px.check_axb(spm, 0.5 * ref.max(), None, b, x) == 1
    # the first column alone passes the backward check at 1e-12; its forward
    # check (threshold 1e4 * eps) passes against the true x0 and fails
    # against an x0 that is off by 1e-6
    b1, x1, x01 = (np.asfortranarray(a[:, :1]) for a in (b, x, x0))
    assert px.check_axb(spm, 1e-12, x01, b1, x1) == 0
    assert px.check_axb(spm, 1e-12, x01 * (1 + 1e-6), b1, x1) == 1


def test_backward_error_propagates_nan():
    spm, A = _random_csc(9, 0.3, "general", False, seed=1)
    b = np.ones((9, 1), order="F")
    x = np.ones((9, 1), order="F")
    x[4, 0] = np.nan
    worst, each = spm.backward_error(b, x)
    assert np.isnan(worst) and np.isnan(each[0])
    assert px.check_axb(spm, 1.0, None, b, x) == 1


def test_rhs_wrappers_refuse_foreign_arrays():
    spm = px.Spm.laplacian7(3, 3, 3)
    n = spm.n
    good = np.zeros((n, 2), order="F")
    px._rhs_array("x", good, spm, -1)
    with pytest.raises(ValueError, match="dtype"):
        px._rhs_array("x", np.zeros((n, 2), dtype=np.float32, order="F"), spm, -1)
    with pytest.raises(ValueError, match="dtype"):
        px._rhs_array("x", np.zeros((n, 2), dtype=np.complex128, order="F"), spm, -1)
    with pytest.raises(ValueError, match="Fortran"):
        px._rhs_array("x", np.zeros((n, 2)), spm, -1)
    with pytest.raises(ValueError, match="rows"):
        px._rhs_array("x", np.zeros((n + 1, 2), order="F"), spm, -1)
    with pytest.raises(ValueError, match="ndarray"):
        px._rhs_array("x", [0.0] * n, spm, -1)
    with pytest.raises(ValueError, match="right-hand sides"):
        px._rhs_array("x", good, spm, 3)
    # task_solve and task_refine check before any call into the library
    with pytest.raises(ValueError):
        px.task_solve(None, spm, np.zeros((n, 2)))
    with pytest.raises(ValueError):
        px.task_refine(None, spm, good, np.zeros((n, 2), dtype=np.float32, order="F"))


def test_matrix_of_another_arithmetic_sets_the_vector_type():
    """Analysis with a real matrix, then spm2bcsc with a complex one of the
    same pattern: the vectors are complex from then on, and applyorder
    permutes complex vectors of the full size both ways."""
    nx = 4
    spm_r = px.Spm.laplacian7(nx, nx, nx)
    spm_c = px.Spm.laplacian7(nx, nx, nx, complex_=True)
    pd, iparm, dparm = analyzed(spm_r, nx)
    assert iparm[px.IPARM_FLOAT] == px.ENUMS["SpmDouble"]
    px._check("spm2bcsc", px.lib.pastix_subtask_spm2bcsc(pd, ctypes.byref(spm_c.s)))
    assert iparm[px.IPARM_FLOAT] == px.ENUMS["SpmComplex64"]
    n = spm_c.n
    rng = np.random.default_rng(0)
    b = np.asfortranarray(rng.uniform(-1, 1, (n, 3)) + 1j * rng.uniform(-1, 1, (n, 3)))
    Bp = px.rhsInit()
    px.subtask_applyorder(pd, px.PastixDirForward, b, Bp)
    perm = px.order_export(pd)[0]
    got = px.rhs_array(Bp)
    assert got.dtype == np.complex128 and got.shape == (n, 3)
    ref = np.zeros_like(b)
    ref[perm] = b
    np.testing.assert_array_equal(got, ref)
    back = np.zeros_like(b)
    px.subtask_applyorder(pd, px.PastixDirBackward, back, Bp)
    np.testing.assert_array_equal(back, b)
    px.rhsFinalize(Bp)
    # and back to a real matrix: real vectors again
    px._check("spm2bcsc", px.lib.pastix_subtask_spm2bcsc(pd, ctypes.byref(spm_r.s)))
    assert iparm[px.IPARM_FLOAT] == px.ENUMS["SpmDouble"]
    Bp = px.rhsInit()
    br = np.asfortranarray(rng.uniform(-1, 1, (n, 2)))
    px.subtask_applyorder(pd, px.PastixDirForward, br, Bp)
    assert px.rhs_array(Bp).dtype == np.float64
    px.rhsFinalize(Bp)
    px.finalize(pd)


def test_steps_out_of_order_are_refused():
    """The reference's order of the steps (sopalin/pastix_task_sopalin.c,
    pastix_task_solve.c): a factorization before the analysis, and a solve or
    a refinement before the factorization, return PASTIX_ERR_BADPARAMETER
    before any work (none reaches the device)."""
    spm = px.Spm.laplacian7(4, 4, 4)
    iparm, dparm = px.initParam()
    iparm[px.IPARM_VERBOSE] = px.PastixVerboseNot
    pd = px.init(iparm, dparm)
    assert px.lib.pastix_task_numfact(pd, ctypes.byref(spm.s)) == px.PASTIX_ERR_BADPARAMETER
    n = spm.n
    x = np.zeros((n, 1), order="F")
    b = np.ones((n, 1), order="F")
    assert px.lib.pastix_task_solve(pd, n, 1, px._ptr(x), n) == px.PASTIX_ERR_BADPARAMETER
    assert px.lib.pastix_task_refine(pd, n, 1, px._ptr(b), n, px._ptr(x), n) == px.PASTIX_ERR_BADPARAMETER
    # after the analysis, still no factorization
    px.task_analyze(pd, spm)
    assert px.lib.pastix_task_solve(pd, n, 1, px._ptr(x), n) == px.PASTIX_ERR_BADPARAMETER
    px.finalize(pd)
