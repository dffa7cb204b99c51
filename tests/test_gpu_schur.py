"""Schur complement API on the device (reference sopalin/schur.c, flow of
example/schur.c): the first third of the unknowns form the Schur part
(pastixSetSchurUnknownList), the factorization leaves them unfactorized
(IPARM_SCHUR_FACT_MODE = PastixFactModeLocal), and

- pastixGetSchur returns S = A22 - A21 A11^{-1} A12, checked against the
  dense complement that numpy computes from the unpermuted matrix (lower
  triangle for Cholesky, whole for LU), in the order the ordering gives the
  Schur unknowns;
- the interface solve of example/schur.c (applyorder, forward sweep on the
  non-Schur part, Schur rows out, dense solve with S, Schur rows back,
  backward sweep, applyorder back) recovers x_true."""
import numpy as np
import pytest
import scipy.linalg as la
import scipy.sparse as sp

import pastix_amd as px
from helpers import make_problem

pytestmark = pytest.mark.gpu


def _dense(spm):
    cp, rp, v = spm.arrays()
    A = sp.csc_matrix((v, rp, cp), shape=(spm.n, spm.n))
    if spm.s.mtxtype == px.SpmSymmetric:
        A = A + A.T - sp.diags(A.diagonal())
    return A.toarray()


@pytest.mark.parametrize("kind,facto", [("lap7", px.PastixFactPOTRF), ("convdiff27", px.PastixFactGETRF)])
def test_schur_complement_and_interface_solve(kind, facto):
    nx = 8
    spm = make_problem(kind, nx)
    n = spm.n
    ns = n // 3
    schur = np.arange(ns, dtype=np.int64)          # baseval 0: the first third
    iparm, dparm = px.initParam()
    iparm[px.IPARM_ORDERING] = px.PastixOrderScotch
    iparm[px.IPARM_FACTORIZATION] = facto
    iparm[px.IPARM_VERBOSE] = px.PastixVerboseNot
    iparm[px.IPARM_SCHUR_FACT_MODE] = px.PastixFactModeLocal
    iparm[px.IPARM_SCHUR_SOLV_MODE] = px.PastixSolvModeInterface
    pd = px.init(iparm, dparm)
    px.setSchurUnknownList(pd, schur)
    px.task_analyze(pd, spm)
    px.task_numfact(pd, spm)
    perm = px.order_export(pd)[0]
    # the Schur unknowns are the last ns of the ordering
    pos = perm[schur] - (n - ns)
    assert np.array_equal(np.sort(pos), np.arange(ns))
    S = np.zeros((ns, ns), order="F")
    px.getSchur(pd, S)
    A = _dense(spm)
    rest = np.setdiff1d(np.arange(n), schur)
    A11, A12 = A[np.ix_(rest, rest)], A[np.ix_(rest, schur)]
    A21, A22 = A[np.ix_(schur, rest)], A[np.ix_(schur, schur)]
    Sref = np.zeros((ns, ns))
    Sref[np.ix_(pos, pos)] = A22 - A21 @ np.linalg.solve(A11, A12)
    lu = facto == px.PastixFactGETRF
    cmp = (lambda X: X) if lu else np.tril
    assert np.abs(cmp(S) - cmp(Sref)).max() <= 1e-12 * np.abs(Sref).max()
    # the interface solve of example/schur.c
    x0, b = spm.gen_rhs(2, seed=4)
    b = np.asfortranarray(b)
    Xp = px.rhsInit()
    px.subtask_applyorder(pd, px.PastixDirForward, b, Xp)
    px.subtask_trsm(pd, px.PastixLeft, px.PastixLower, px.PastixNoTrans,
                    px.PastixUnit if lu else px.PastixNonUnit, Xp)
    xs = np.zeros((ns, 2), order="F")
    px.rhsSchurGet(pd, Xp, xs)
    # the Schur system in the ordering's numbering of the Schur block, as
    # pastixGetSchur and pastixRhsSchurGet give it
    if lu:
        xs = la.lu_solve(la.lu_factor(S), xs)
    else:
        xs = la.cho_solve(la.cho_factor(S, lower=True), xs)
    px.rhsSchurSet(pd, xs, Xp)
    if lu:
        px.subtask_trsm(pd, px.PastixLeft, px.PastixUpper, px.PastixNoTrans, px.PastixNonUnit, Xp)
    else:
        px.subtask_trsm(pd, px.PastixLeft, px.PastixLower, px.PastixConjTrans, px.PastixNonUnit, Xp)
    x = np.zeros_like(b)
    px.subtask_applyorder(pd, px.PastixDirBackward, x, Xp)
    px.rhsFinalize(Xp)
    assert np.abs(x - x0).max() <= 1e-10 * np.abs(x0).max()
    px.finalize(pd)


def test_get_schur_without_a_list_is_refused():
    nx = 5
    spm = make_problem("lap7", nx)
    iparm, dparm = px.initParam()
    iparm[px.IPARM_ORDERING] = px.PastixOrderScotch
    iparm[px.IPARM_VERBOSE] = px.PastixVerboseNot
    pd = px.init(iparm, dparm)
    px.task_analyze(pd, spm)
    px.task_numfact(pd, spm)
    S = np.zeros((4, 4), order="F")
    with pytest.raises(px.PastixError) as e:
        px.getSchur(pd, S)
    assert e.value.code == px.PASTIX_ERR_BADPARAMETER
    px.finalize(pd)
