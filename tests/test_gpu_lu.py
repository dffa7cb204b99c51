"""LU with static pivoting on the MI355X engine (GETRF path) against the CPU
oracle: every L and U^T tile (cpucblk_zdiff criterion, reference
kernels/cpucblk_zdiff.c:93-100), the solve, and FP32 LU + FP64 GMRES on the
27-point convection-diffusion operator (unsymmetric values, symmetric
pattern)."""
import numpy as np
import pytest

import pastix_amd as px
import oracle_ref as orc
from helpers import analyzed, cblk_diff, make_problem, static_pivot_threshold

pytestmark = pytest.mark.gpu


def _lu_case(nx, mixed=0):
    spm = make_problem("convdiff27", nx, seed=1)
    pd, iparm, dparm = analyzed(spm, nx, facto=px.PastixFactLU, mixed=mixed)
    px.task_numfact(pd, spm)
    return spm, pd, iparm, dparm


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("nx", [4, 8, 13])
def test_getrf_tiles_match_oracle(nx, dtype):
    spm, pd, iparm, dparm = _lu_case(nx, mixed=int(dtype == np.float32))
    s = px.solver_export(pd)
    L = px.get_coeftab(pd, s["coefnbr"], dtype, side=0)
    U = px.get_coeftab(pd, s["coefnbr"], dtype, side=1)
    F = orc.OracleLU(s, px.bcsc_export(pd), static_pivot_threshold(dparm, dtype), dtype)
    assert iparm[px.IPARM_STATIC_PIVOTING] == F.npiv == 0
    assert cblk_diff(s, L, F.L, dtype, full_diag=True) <= 10.0
    assert cblk_diff(s, U, F.U, dtype, full_diag=True) <= 10.0
    px.finalize(pd)


@pytest.mark.parametrize("nx,nrhs", [(8, 1), (14, 3)])
def test_lu_solve_backward_error(nx, nrhs):
    spm, pd, iparm, dparm = _lu_case(nx)
    x0, b = spm.gen_rhs(nrhs, seed=0)
    x = np.array(b, order="F", copy=True)
    px.task_solve(pd, spm, x)
    berr, _ = spm.backward_error(b, x)
    assert berr <= 1e-13
    px.finalize(pd)


def test_lu_mixed_gmres_reaches_1e12():
    spm, pd, iparm, dparm = _lu_case(12, mixed=1)
    x0, b = spm.gen_rhs(2, seed=4)
    x = np.array(b, order="F", copy=True)
    px.task_solve(pd, spm, x)
    px.set_option(pd, "berr_target", 1e-12)
    px.task_refine(pd, spm, b, x)
    berr, _ = spm.backward_error(b, x)
    assert berr <= 1e-12
    assert 0 < iparm[px.IPARM_NBITER] < 30
    px.finalize(pd)
