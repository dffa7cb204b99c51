"""Parity of the MI355X engine with the CPU oracle (GPU tests).

The device factor is compared tile by tile with the oracle's on the same
solver structure (cpucblk_zdiff criterion, reference kernels/cpucblk_zdiff.c:
93-100), and the solutions through the C API are checked by backward error
(configuration 1 asks 1e-13 for the 20^3 Laplacian).
"""
import numpy as np
import pytest

import pastix_amd as px
import oracle_ref as orc
from helpers import analyzed, cblk_diff, make_problem, permute_rhs, static_pivot_threshold

pytestmark = pytest.mark.gpu


def _factor_case(nx, mixed=0, kernels=1):
    spm = make_problem("lap7", nx)
    pd, iparm, dparm = analyzed(spm, nx, mixed=mixed)
    px.set_option(pd, "kernels", kernels)
    px.task_numfact(pd, spm)
    return spm, pd, iparm, dparm


@pytest.mark.parametrize("kernels", [1, 0])
@pytest.mark.parametrize("nx", [3, 5, 8, 12, 20, 28])
def test_potrf_fp64_tiles_match_oracle(nx, kernels):
    spm, pd, iparm, dparm = _factor_case(nx, kernels=kernels)
    s = px.solver_export(pd)
    L = px.get_coeftab(pd, s["coefnbr"], np.float64)
    F = orc.OracleFactor(s, px.bcsc_export(pd), static_pivot_threshold(dparm, np.float64))
    assert iparm[px.IPARM_STATIC_PIVOTING] == F.npiv == 0
    assert cblk_diff(s, L, F.L, np.float64) <= 10.0
    px.finalize(pd)


@pytest.mark.parametrize("kernels", [1, 0])
@pytest.mark.parametrize("nx", [5, 12, 24])
def test_potrf_fp32_tiles_match_oracle(nx, kernels):
    spm, pd, iparm, dparm = _factor_case(nx, mixed=1, kernels=kernels)
    s = px.solver_export(pd)
    L = px.get_coeftab(pd, s["coefnbr"], np.float32)
    F = orc.OracleFactor(s, px.bcsc_export(pd), static_pivot_threshold(dparm, np.float32), np.float32)
    assert cblk_diff(s, L, F.L, np.float32) <= 10.0
    px.finalize(pd)


@pytest.mark.parametrize("kernels", [1, 0])
@pytest.mark.parametrize("nx,nrhs", [(8, 1), (20, 1), (20, 3), (12, 9), (30, 2)])
def test_solve_fp64_backward_error(nx, nrhs, kernels):
    spm, pd, iparm, dparm = _factor_case(nx, kernels=kernels)
    x0, b = spm.gen_rhs(nrhs, seed=0)
    x = np.array(b, order="F", copy=True)
    px.task_solve(pd, spm, x)
    berr, _ = spm.backward_error(b, x)
    assert berr <= 1e-13
    assert np.abs(x - x0).max() <= 1e-10
    px.finalize(pd)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("big_rows,timing,panel_mw,small", [(1, 0, 1, 0), (10 ** 9, 0, 1, 0), (10 ** 9, 0, 1, 1),
                                                           (192, 1, 1, 1), (192, 0, 0, 1)])
def test_update_tilings_and_schedules_match_oracle(dtype, big_rows, timing, panel_mw, small):
    """Every blok update on the tall tiles (big_rows=1), every one on the
    64 x 64 tiles (big_rows=1e9), the small updates on one wavefront each
    (small=1) or on tiles, the single-stream timed schedule instead of the
    two-stream lookahead, and the single-workgroup panel instead of the
    one-workgroup-per-block panel: the factor matches the oracle."""
    nx = 18
    mixed = int(dtype == np.float32)
    spm = make_problem("lap7", nx)
    pd, iparm, dparm = analyzed(spm, nx, mixed=mixed)
    px.set_option(pd, "upd_big_rows", big_rows)
    px.set_option(pd, "timing", timing)
    px.set_option(pd, "panel_mw", panel_mw)
    px.set_option(pd, "upd_small", small)
    px.task_numfact(pd, spm)
    s = px.solver_export(pd)
    L = px.get_coeftab(pd, s["coefnbr"], dtype)
    F = orc.OracleFactor(s, px.bcsc_export(pd), static_pivot_threshold(dparm, dtype), dtype)
    assert cblk_diff(s, L, F.L, dtype) <= 10.0
    px.finalize(pd)


@pytest.mark.parametrize("dma,pad", [(1, 4096), (1, 0), (0, 0)])
@pytest.mark.parametrize("kind,nx", [("lap7", 20), ("convdiff27", 13)])
def test_single_tall_tiles_match_oracle(kind, nx, dma, pad):
    """Every blok update of a single-precision factor on the 256-row tiles
    (upd_big_rows=1), through the kernel that loads its operands straight
    into LDS (upd_dma=1, with and without the LDS padding that limits the
    workgroups per CU) or through the register-staged one (upd_dma=0), for
    LL^T and for LU (the U-side tiles too): L (and U) match the oracle.  The 20^3 separators give cblks wider
    than two stages and facing bloks narrower than the 64-column tile."""
    lu = kind == "convdiff27"
    spm = make_problem(kind, nx)
    pd, iparm, dparm = analyzed(spm, nx, facto=px.PastixFactGETRF if lu else None, mixed=1)
    px.set_option(pd, "upd_big_rows", 1)
    px.set_option(pd, "upd_small", 0)
    px.set_option(pd, "upd_dma", dma)
    px.set_option(pd, "upd_dma_pad", pad)
    px.task_numfact(pd, spm)
    s = px.solver_export(pd)
    L = px.get_coeftab(pd, s["coefnbr"], np.float32)
    if lu:
        F = orc.OracleLU(s, px.bcsc_export(pd), static_pivot_threshold(dparm, np.float32), np.float32)
        U = px.get_coeftab(pd, s["coefnbr"], np.float32, side=1)
        assert cblk_diff(s, L, F.L, np.float32, full_diag=True) <= 10.0
        assert cblk_diff(s, U, F.U, np.float32, full_diag=True) <= 10.0
    else:
        F = orc.OracleFactor(s, px.bcsc_export(pd), static_pivot_threshold(dparm, np.float32), np.float32)
        assert cblk_diff(s, L, F.L, np.float32) <= 10.0
    px.finalize(pd)


def test_tall_tile_kernel_option_values():
    spm = make_problem("lap7", 6)
    pd, iparm, dparm = analyzed(spm, 6, mixed=1)
    with pytest.raises(px.PastixError):
        px.set_option(pd, "upd_dma_pad", 1 << 20)
    px.finalize(pd)


@pytest.mark.parametrize("kind,facto,dtype", [("lap7", None, np.float64), ("lap7", None, np.float32),
                                               ("convdiff27", px.PastixFactGETRF, np.float64),
                                               ("convdiff27", px.PastixFactGETRF, np.float32),
                                               ("lap7", px.PastixFactLDLT, np.float64)])
def test_one_wavefront_updates_over_their_whole_range_match_oracle(kind, facto, dtype):
    """Every blok update of at most 256 stacked rows, 64 facing columns and
    32 source columns on the one-wavefront kernel (one wavefront per 64
    rows, the facing columns by chunks of 16): the factor matches the
    oracle (L and U for LU; LDL^T with its D scaling)."""
    nx = 14
    spm = make_problem(kind, nx)
    mixed = int(dtype == np.float32)
    pd, iparm, dparm = analyzed(spm, nx, facto=facto, mixed=mixed)
    px.set_option(pd, "upd_small", 1)
    px.set_option(pd, "upd_small_m", 256)
    px.set_option(pd, "upd_small_n", 64)
    px.task_numfact(pd, spm)
    s = px.solver_export(pd)
    L = px.get_coeftab(pd, s["coefnbr"], dtype)
    tau = static_pivot_threshold(dparm, dtype)
    if facto == px.PastixFactGETRF:
        F = orc.OracleLU(s, px.bcsc_export(pd), tau, dtype)
        U = px.get_coeftab(pd, s["coefnbr"], dtype, side=1)
        assert cblk_diff(s, L, F.L, dtype, full_diag=True) <= 10.0
        assert cblk_diff(s, U, F.U, dtype, full_diag=True) <= 10.0
    elif facto == px.PastixFactLDLT:
        F = orc.OracleLDLT(s, px.bcsc_export(pd), tau, dtype)
        assert cblk_diff(s, L, F.L, dtype, full_diag=True) <= 10.0
    else:
        F = orc.OracleFactor(s, px.bcsc_export(pd), tau, dtype)
        assert cblk_diff(s, L, F.L, dtype) <= 10.0
    px.finalize(pd)


def test_device_solve_matches_oracle_solve():
    nx = 12
    spm, pd, iparm, dparm = _factor_case(nx)
    s = px.solver_export(pd)
    F = orc.OracleFactor(s, px.bcsc_export(pd), static_pivot_threshold(dparm, np.float64))
    x0, b = spm.gen_rhs(2, seed=3)
    bp = permute_rhs(pd, b)
    xo = F.solve(bp)
    x = np.array(b, order="F", copy=True)
    px.task_solve(pd, spm, x)
    xp = permute_rhs(pd, x)
    assert np.abs(xp - xo).max() <= 1e-12 * max(1.0, np.abs(xo).max())
    px.finalize(pd)


@pytest.mark.parametrize("spread", [1, 2, 0])
@pytest.mark.parametrize("facto,mixed,nrhs", [(None, 0, 2), (None, 1, 4), (px.PastixFactGETRF, 0, 3),
                                              (px.PastixFactLDLT, 0, 7)])
def test_backward_partial_sums_match_oracle_solve(spread, facto, mixed, nrhs):
    """The backward sweep's gather into partial-sum copies (solve_spread = 1:
    every cblk with an off-diagonal tile, 2: those with two or more, 0: the
    direct atomics): the solve matches the oracle's, for LL^T, LU and LDL^T,
    single and double factors, 2 to 7 right-hand sides (7 takes the 8-column
    launches, where the double factor keeps the direct atomics)."""
    nx = 14
    kind = "convdiff27" if facto == px.PastixFactGETRF else "lap7"
    spm = make_problem(kind, nx)
    pd, iparm, dparm = analyzed(spm, nx, facto=facto, mixed=mixed)
    px.set_option(pd, "solve_spread", spread)
    px.task_numfact(pd, spm)
    dtype = np.float32 if mixed else np.float64
    s = px.solver_export(pd)
    tau = static_pivot_threshold(dparm, dtype)
    Or = {None: orc.OracleFactor, px.PastixFactGETRF: orc.OracleLU, px.PastixFactLDLT: orc.OracleLDLT}[facto]
    F = Or(s, px.bcsc_export(pd), tau, dtype)
    x0, b = spm.gen_rhs(nrhs, seed=5)
    bp = permute_rhs(pd, b)
    xo = F.solve(bp)
    for rep in range(2):   # the copies are cleared again for the second solve
        x = np.array(b, order="F", copy=True)
        px.task_solve(pd, spm, x)
        xp = permute_rhs(pd, x)
        tol = (1e-4 if mixed else 1e-12) * max(1.0, np.abs(xo).max())
        assert np.abs(xp - xo).max() <= tol, rep
    px.finalize(pd)


def test_backward_partial_sums_option_values():
    spm = make_problem("lap7", 6)
    pd, iparm, dparm = analyzed(spm, 6)
    with pytest.raises(px.PastixError):
        px.set_option(pd, "solve_spread", -1)
    px.finalize(pd)


@pytest.mark.parametrize("nx,nrhs", [(8, 1), (16, 4)])
def test_mixed_fp32_gmres_reaches_1e12(nx, nrhs):
    spm, pd, iparm, dparm = _factor_case(nx, mixed=1)
    x0, b = spm.gen_rhs(nrhs, seed=0)
    x = np.array(b, order="F", copy=True)
    px.task_solve(pd, spm, x)
    berr_solve, _ = spm.backward_error(b, x)
    assert berr_solve > 1e-12          # a single precision solve alone is not enough
    px.set_option(pd, "berr_target", 1e-12)
    px.task_refine(pd, spm, b, x)
    berr, _ = spm.backward_error(b, x)
    assert berr <= 1e-12
    assert 0 < iparm[px.IPARM_NBITER] < 20
    px.finalize(pd)


def test_fp64_gmres_reference_criterion():
    nx = 10
    spm, pd, iparm, dparm = _factor_case(nx)
    x0, b = spm.gen_rhs(1, seed=1)
    x = np.zeros_like(b, order="F")      # start from zero: GMRES has work to do
    px.task_refine(pd, spm, b, x)
    r = b - spm.matmat(x)
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-12
    px.finalize(pd)


@pytest.mark.parametrize("kind,facto,mixed,nx", [("lap7", None, 0, 12), ("lap7", None, 1, 14),
                                                 ("convdiff27", px.PastixFactGETRF, 0, 8),
                                                 ("lap7sd", px.PastixFactLDLT, 0, 9)])
def test_factorization_graph_replays_match_oracle(kind, facto, mixed, nx):
    """The level loop of the factorization captured as a graph (engine
    option fact_graph = 1, the default) and replayed: every replay, with
    solves and an eager factorization (fact_graph = 0) in between, gives a
    factor that matches the oracle's tile by tile."""
    spm = make_problem(kind, nx)
    pd, iparm, dparm = analyzed(spm, nx, facto=facto, mixed=mixed)
    dt = np.float32 if mixed else np.float64
    for rep, graph in enumerate((1, 1, 0, 1, 1)):
        px.set_option(pd, "fact_graph", graph)
        px.task_numfact(pd, spm)
        if rep == 0:
            s = px.solver_export(pd)
            bc = px.bcsc_export(pd)
            tau = static_pivot_threshold(dparm, dt)
            if facto == px.PastixFactGETRF:
                F = orc.OracleLU(s, bc, tau, dt)
            elif facto == px.PastixFactLDLT:
                F = orc.OracleLDLT(s, bc, tau, dt)
            else:
                F = orc.OracleFactor(s, bc, tau, dt)
        L = px.get_coeftab(pd, s["coefnbr"], dt)
        assert cblk_diff(s, L, F.L, dt, full_diag=facto is not None) <= 10.0
        if facto == px.PastixFactGETRF:
            U = px.get_coeftab(pd, s["coefnbr"], dt, side=1)
            assert cblk_diff(s, U, F.U, dt, full_diag=True) <= 10.0
        x0, b = spm.gen_rhs(2, seed=rep)
        x = np.array(b, order="F", copy=True)
        px.task_solve(pd, spm, x)
    px.finalize(pd)
