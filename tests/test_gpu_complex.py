"""Complex arithmetic on the MI355X engine against the CPU oracle: the
complex-symmetric LDL^T (A = L7 - (0.5 + 0.1i) I, plain transposes, reference
kernels/core_zsytrfsp.c) and the complex LU, in complex128 and complex64,
tile by tile (cpucblk_zdiff criterion), the solve, and complex64 LDL^T +
complex128 BiCGStab to a backward error of 1e-12."""
import numpy as np
import pytest

import pastix_amd as px
import oracle_ref as orc
from helpers import analyzed, cblk_diff, static_pivot_threshold

pytestmark = pytest.mark.gpu


def _case(nx, facto, mixed=0, extra=None):
    spm = px.Spm.helmholtz7(nx, nx, nx, 0.5, 0.1)
    pd, iparm, dparm = analyzed(spm, nx, facto=facto, mixed=mixed, extra=extra)
    px.task_numfact(pd, spm)
    return spm, pd, iparm, dparm


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("nx", [5, 9, 13])
def test_complex_sytrf_tiles_match_oracle(nx, dtype):
    spm, pd, iparm, dparm = _case(nx, px.PastixFactLDLT, mixed=int(dtype == np.complex64))
    s = px.solver_export(pd)
    L = px.get_coeftab(pd, s["coefnbr"], dtype)
    F = orc.OracleLDLT(s, px.bcsc_export(pd, complex_=True), static_pivot_threshold(dparm, dtype), dtype)
    assert iparm[px.IPARM_STATIC_PIVOTING] == F.npiv == 0
    assert cblk_diff(s, L, F.L, dtype, full_diag=True) <= 10.0
    px.finalize(pd)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_complex_getrf_tiles_match_oracle(dtype):
    spm, pd, iparm, dparm = _case(9, px.PastixFactLU, mixed=int(dtype == np.complex64))
    s = px.solver_export(pd)
    L = px.get_coeftab(pd, s["coefnbr"], dtype, side=0)
    U = px.get_coeftab(pd, s["coefnbr"], dtype, side=1)
    F = orc.OracleLU(s, px.bcsc_export(pd, complex_=True), static_pivot_threshold(dparm, dtype), dtype)
    assert cblk_diff(s, L, F.L, dtype, full_diag=True) <= 10.0
    assert cblk_diff(s, U, F.U, dtype, full_diag=True) <= 10.0
    px.finalize(pd)


@pytest.mark.parametrize("nx,nrhs", [(8, 1), (12, 3)])
def test_complex_ldlt_solve_backward_error(nx, nrhs):
    spm, pd, iparm, dparm = _case(nx, px.PastixFactLDLT)
    x0, b = spm.gen_rhs(nrhs, seed=0)
    x = np.array(b, order="F", copy=True)
    px.task_solve(pd, spm, x)
    berr, _ = spm.backward_error(b, x)
    assert berr <= 1e-13
    px.finalize(pd)


def test_complex64_ldlt_bicgstab_reaches_1e12():
    spm, pd, iparm, dparm = _case(12, px.PastixFactLDLT, mixed=1,
                                  extra={px.IPARM_REFINEMENT: px.PastixRefineBiCGSTAB})
    x0, b = spm.gen_rhs(2, seed=1)
    x = np.array(b, order="F", copy=True)
    px.task_solve(pd, spm, x)
    px.set_option(pd, "berr_target", 1e-12)
    px.task_refine(pd, spm, b, x)
    berr, _ = spm.backward_error(b, x)
    assert berr <= 1e-12
    assert 0 < iparm[px.IPARM_NBITER] < 30
    px.finalize(pd)
