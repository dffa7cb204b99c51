"""The one-launch solve sweeps (engine option solve_sf = 1, the default;
passes of at most 4 right-hand sides) against the level-by-level sweeps
(solve_sf = 0) and the oracle's solve with the same factor, for every
factorization and arithmetic, with right-hand side counts that take one
pass and several (20 RHS: passes of 8, level by level).

Both device solves compute the same sums in different orders (atomic
accumulation), so they agree to rounding of the factor precision; the
oracle's supernodal solve (reference kernels/solve_ztrsmsp.c) with the
device's own factor bounds both."""
import copy

import numpy as np
import pytest

import pastix_amd as px
import oracle_ref as orc
from helpers import analyzed, make_problem, permute_rhs, static_pivot_threshold, unpermute

pytestmark = pytest.mark.gpu

CASES = [
    # (matrix, facto, mixed, grid edge)
    ("lap7", None, 0, 11),
    ("lap7", None, 1, 11),
    ("lap7sd", px.PastixFactLDLT, 0, 10),
    ("convdiff27", px.PastixFactGETRF, 0, 9),
    ("convdiff27", px.PastixFactGETRF, 1, 9),
    ("helm", px.PastixFactLDLT, 0, 9),
    ("helm", px.PastixFactLDLT, 1, 9),
    ("helm", px.PastixFactGETRF, 0, 8),
]


def _problem(kind, nx):
    if kind == "helm":
        return px.Spm.helmholtz7(nx, nx, nx, 0.5, 0.1)
    return make_problem(kind, nx)


@pytest.mark.parametrize("cut", [0, 3, -1])
@pytest.mark.parametrize("nrhs", [1, 4, 20])
@pytest.mark.parametrize("kind,facto,mixed,nx", CASES)
def test_one_launch_solve_matches_level_solve(kind, facto, mixed, nx, nrhs, cut):
    """cut: the first level of the one-launch part (0: every level, 3: the
    levels below run level by level, -1: the automatic cut)."""
    spm = _problem(kind, nx)
    pd, iparm, dparm = analyzed(spm, nx, facto=facto, mixed=mixed)
    px.set_option(pd, "sf_from_level", cut)
    px.task_numfact(pd, spm)
    x0, b = spm.gen_rhs(nrhs, seed=3)
    xs = []
    for sf in (1, 0, 1):
        px.set_option(pd, "solve_sf", sf)
        x = np.array(b, order="F", copy=True)
        px.task_solve(pd, spm, x)
        xs.append(x)
    single = mixed == 1
    eps = np.finfo(np.float32 if single else np.float64).eps
    scale = np.abs(xs[1]).max()
    # one-launch vs level sweeps, and one-launch run twice (the counters
    # carry over from pass to pass)
    assert np.abs(xs[0] - xs[1]).max() <= 200 * eps * scale
    assert np.abs(xs[2] - xs[1]).max() <= 200 * eps * scale
    berr_sf = spm.backward_error(b, xs[0])[0]
    berr_lv = spm.backward_error(b, xs[1])[0]
    assert berr_sf <= 10 * max(berr_lv, eps)
    if facto in (None, px.PastixFactLDLT) and not single:
        # the oracle's solve with the device factor
        s = px.solver_export(pd)
        cplx = spm.is_complex
        dt = np.complex128 if cplx else np.float64
        L = px.get_coeftab(pd, s["coefnbr"], dt)
        bc = px.bcsc_export(pd, complex_=cplx)
        tau = static_pivot_threshold(dparm, dt)
        F = (orc.OracleLDLT if facto == px.PastixFactLDLT else orc.OracleFactor)(s, bc, tau, dt)
        G = copy.copy(F)
        G.L = L
        xo = unpermute(pd, G.solve(permute_rhs(pd, b)))
        assert np.abs(xs[0] - xo).max() <= 200 * eps * scale
    px.finalize(pd)


def test_one_launch_solve_on_a_refactorized_instance():
    """A new factorization (same structure) and a structure change keep the
    counters and task lists consistent."""
    nx = 10
    spm = make_problem("lap7", nx)
    pd, iparm, dparm = analyzed(spm, nx)
    for rep in range(2):
        px.task_numfact(pd, spm)
        x0, b = spm.gen_rhs(3, seed=rep)
        x = np.array(b, order="F", copy=True)
        px.task_solve(pd, spm, x)
        assert spm.backward_error(b, x)[0] <= 1e-14
    px.finalize(pd)
    spm2 = make_problem("lap7", nx + 3)
    pd2, _, _ = analyzed(spm2, nx + 3)
    px.task_numfact(pd2, spm2)
    x0, b = spm2.gen_rhs(2, seed=5)
    x = np.array(b, order="F", copy=True)
    px.task_solve(pd2, spm2, x)
    assert spm2.backward_error(b, x)[0] <= 1e-14
    px.finalize(pd2)


@pytest.mark.parametrize("kind,facto,nx", [("lap7", None, 12), ("convdiff27", px.PastixFactGETRF, 9),
                                           ("helm", px.PastixFactLDLT, 9), ("helm", px.PastixFactGETRF, 8)])
def test_single_factor_on_double_vectors(kind, facto, nx):
    """A single-precision factor applied to double vectors (option
    solve_vec_double = 1, the entries widened as they are read) against the
    reference's way (vectors rounded to single and back): the same solution
    to single-precision accuracy, a backward error no worse, and a mixed
    GMRES / BiCGStab that needs no more iterations."""
    spm = _problem(kind, nx)
    extra = {px.IPARM_REFINEMENT: px.PastixRefineBiCGSTAB} if kind == "helm" else None
    pd, iparm, dparm = analyzed(spm, nx, facto=facto, mixed=1, extra=extra)
    px.task_numfact(pd, spm)
    x0, b = spm.gen_rhs(3, seed=7)
    xs, berr, iters = {}, {}, {}
    for vd in (1, 0):
        px.set_option(pd, "solve_vec_double", vd)
        x = np.array(b, order="F", copy=True)
        px.task_solve(pd, spm, x)
        xs[vd] = x
        berr[vd] = spm.backward_error(b, x)[0]
        px.set_option(pd, "berr_target", 1e-12)
        xr = np.array(x, order="F", copy=True)
        px.task_refine(pd, spm, b, xr)
        assert spm.backward_error(b, xr)[0] <= 1e-12
        iters[vd] = iparm[px.IPARM_NBITER]
    eps = np.finfo(np.float32).eps
    assert np.abs(xs[1] - xs[0]).max() <= 1e3 * eps * np.abs(xs[0]).max()
    assert berr[1] <= 1.5 * berr[0]
    assert iters[1] <= iters[0]
    px.finalize(pd)
