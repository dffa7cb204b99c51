"""LDL^T with static pivoting on the MI355X engine (SYTRF path) against the
CPU oracle: every tile including the D L^T upper part of the diagonal tiles
(cpucblk_zdiff criterion, reference kernels/cpucblk_zdiff.c:93-100) on a
diagonally dominant indefinite matrix (no element growth, so two correct
implementations agree to rounding), then the solve and FP32 LDL^T + FP64
GMRES on the 7-point Laplacian shifted to an indefinite spectrum (diagonal
4.3)."""
import numpy as np
import pytest

import pastix_amd as px
import oracle_ref as orc
from helpers import analyzed, cblk_diff, make_problem, permute_rhs, static_pivot_threshold, unpermute

pytestmark = pytest.mark.gpu


def _ldlt_case(nx, mixed=0, kind="lap7"):
    spm = make_problem(kind, nx, diag=4.3) if kind == "lap7" else make_problem(kind, nx)
    pd, iparm, dparm = analyzed(spm, nx, facto=px.PastixFactLDLT, mixed=mixed)
    px.task_numfact(pd, spm)
    return spm, pd, iparm, dparm


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("nx", [5, 9, 14])
def test_sytrf_tiles_match_oracle(nx, dtype):
    spm, pd, iparm, dparm = _ldlt_case(nx, mixed=int(dtype == np.float32), kind="lap7sd")
    s = px.solver_export(pd)
    L = px.get_coeftab(pd, s["coefnbr"], dtype)
    F = orc.OracleLDLT(s, px.bcsc_export(pd), static_pivot_threshold(dparm, dtype), dtype)
    assert iparm[px.IPARM_STATIC_PIVOTING] == F.npiv
    assert cblk_diff(s, L, F.L, dtype, full_diag=True) <= 10.0
    px.finalize(pd)


@pytest.mark.parametrize("nx,nrhs", [(8, 1), (13, 3)])
def test_ldlt_solve_backward_error(nx, nrhs):
    spm, pd, iparm, dparm = _ldlt_case(nx)
    x0, b = spm.gen_rhs(nrhs, seed=0)
    x = np.array(b, order="F", copy=True)
    px.task_solve(pd, spm, x)
    berr, _ = spm.backward_error(b, x)
    # LDL^T without pivoting of an indefinite matrix has element growth, so
    # the bound is the oracle's own solve on the same structure, within 10x
    s = px.solver_export(pd)
    F = orc.OracleLDLT(s, px.bcsc_export(pd), static_pivot_threshold(dparm, np.float64), np.float64)
    xo = unpermute(pd, F.solve(permute_rhs(pd, b)))
    berr_o, _ = spm.backward_error(b, np.asfortranarray(xo))
    assert berr <= max(1e-13, 10.0 * berr_o)
    px.finalize(pd)


def test_ldlt_mixed_gmres_reaches_1e12():
    spm, pd, iparm, dparm = _ldlt_case(12, mixed=1)
    x0, b = spm.gen_rhs(2, seed=6)
    x = np.array(b, order="F", copy=True)
    px.task_solve(pd, spm, x)
    px.set_option(pd, "berr_target", 1e-12)
    px.task_refine(pd, spm, b, x)
    berr, _ = spm.backward_error(b, x)
    assert berr <= 1e-12
    assert 0 < iparm[px.IPARM_NBITER] < 30
    px.finalize(pd)
