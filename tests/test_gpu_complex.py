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


@pytest.mark.parametrize("facto", [px.PastixFactLDLT, px.PastixFactLU])
@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_complex_tall_tile_kernel_matches_oracle(dtype, facto):
    """Every update that is not a one-wavefront update goes to the tall-tile
    kernel (upd_big_rows = 1), whose complex loads move both halves of each
    value in one buffer load."""
    nx = 9
    spm = px.Spm.helmholtz7(nx, nx, nx, 0.5, 0.1)
    pd, iparm, dparm = analyzed(spm, nx, facto=facto, mixed=int(dtype == np.complex64))
    px.set_option(pd, "upd_big_rows", 1)
    px.task_numfact(pd, spm)
    s = px.solver_export(pd)
    bc = px.bcsc_export(pd, complex_=True)
    tau = static_pivot_threshold(dparm, dtype)
    L = px.get_coeftab(pd, s["coefnbr"], dtype, side=0)
    if facto == px.PastixFactLDLT:
        F = orc.OracleLDLT(s, bc, tau, dtype)
    else:
        F = orc.OracleLU(s, bc, tau, dtype)
        U = px.get_coeftab(pd, s["coefnbr"], dtype, side=1)
        assert cblk_diff(s, U, F.U, dtype, full_diag=True) <= 10.0
    assert cblk_diff(s, L, F.L, dtype, full_diag=True) <= 10.0
    px.finalize(pd)


@pytest.mark.parametrize("facto", [px.PastixFactLDLT, px.PastixFactLU])
@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_complex_one_wavefront_updates_over_their_whole_range(dtype, facto):
    """Every update of at most 256 stacked rows, 64 facing columns and 32
    source columns on the one-wavefront kernel (upd_small_m = 256,
    upd_small_n = 64)."""
    nx = 11
    spm = px.Spm.helmholtz7(nx, nx, nx, 0.5, 0.1)
    pd, iparm, dparm = analyzed(spm, nx, facto=facto, mixed=int(dtype == np.complex64))
    px.set_option(pd, "upd_small_m", 256)
    px.set_option(pd, "upd_small_n", 64)
    px.task_numfact(pd, spm)
    s = px.solver_export(pd)
    bc = px.bcsc_export(pd, complex_=True)
    tau = static_pivot_threshold(dparm, dtype)
    L = px.get_coeftab(pd, s["coefnbr"], dtype, side=0)
    if facto == px.PastixFactLDLT:
        F = orc.OracleLDLT(s, bc, tau, dtype)
    else:
        F = orc.OracleLU(s, bc, tau, dtype)
        U = px.get_coeftab(pd, s["coefnbr"], dtype, side=1)
        assert cblk_diff(s, U, F.U, dtype, full_diag=True) <= 10.0
    assert cblk_diff(s, L, F.L, dtype, full_diag=True) <= 10.0
    px.finalize(pd)


@pytest.mark.parametrize("facto", [px.PastixFactLDLT, px.PastixFactLU])
def test_complex128_cblk_wider_than_lds_rows(facto):
    """17^3: the top separator is one 289-column cblk.  In complex double the
    panel's 257-row L21 strip and the inverse's 257-row R block exceed the
    rows the CU's LDS holds next to the kernels' static tiles (about 244), so
    the rows beyond are kept in HBM; the factor must not change."""
    nx = 17
    spm, pd, iparm, dparm = _case(nx, facto)
    s = px.solver_export(pd)
    cb = s["cblk"].reshape(-1, 6)[:s["cblknbr"]]
    assert (cb[:, 1] - cb[:, 0] + 1).max() == 289
    bc = px.bcsc_export(pd, complex_=True)
    tau = static_pivot_threshold(dparm, np.complex128)
    L = px.get_coeftab(pd, s["coefnbr"], np.complex128, side=0)
    if facto == px.PastixFactLDLT:
        F = orc.OracleLDLT(s, bc, tau, np.complex128)
        assert cblk_diff(s, L, F.L, np.complex128, full_diag=True) <= 10.0
    else:
        F = orc.OracleLU(s, bc, tau, np.complex128)
        U = px.get_coeftab(pd, s["coefnbr"], np.complex128, side=1)
        assert cblk_diff(s, L, F.L, np.complex128, full_diag=True) <= 10.0
        assert cblk_diff(s, U, F.U, np.complex128, full_diag=True) <= 10.0
    x0, b = spm.gen_rhs(2, seed=0)
    x = np.array(b, order="F", copy=True)
    px.task_solve(pd, spm, x)
    berr, _ = spm.backward_error(b, x)
    assert berr <= 1e-13
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
