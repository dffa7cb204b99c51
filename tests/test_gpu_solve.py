"""The level-by-level solve sweeps against the oracle's solve with the
device's own factor, for every factorization, arithmetic and precision
(double and single factors, real and complex; Cholesky, LDL^T and LU), with
right-hand side counts that take one pass and several (20 RHS: passes of 8),
with the passes replayed from captured graphs (engine option solve_graph =
1, the default) and launched kernel by kernel (0); and the graph-replayed
solve against the same sweeps launched one by one, replay after replay,
with other device work in between.

The device and the oracle's supernodal solve (reference
kernels/solve_ztrsmsp.c) compute the same sums in different orders (atomic
accumulation on the device), so they agree to rounding of the factor
precision."""
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


@pytest.mark.parametrize("graph", [1, 0])
@pytest.mark.parametrize("nrhs", [1, 4, 20])
@pytest.mark.parametrize("kind,facto,mixed,nx", CASES)
def test_level_solve_matches_oracle_solve(kind, facto, mixed, nx, nrhs, graph):
    """The device solve (graph-replayed passes or launched kernel by kernel)
    against the oracle's supernodal solve with the device's own factor, in
    the factor precision (a single factor: vectors rounded to single, as the
    device's mixed solve does); both compute the same sums in other orders."""
    spm = _problem(kind, nx)
    pd, iparm, dparm = analyzed(spm, nx, facto=facto, mixed=mixed)
    px.set_option(pd, "solve_graph", graph)
    px.task_numfact(pd, spm)
    x0, b = spm.gen_rhs(nrhs, seed=3)
    x = np.array(b, order="F", copy=True)
    px.task_solve(pd, spm, x)
    single = mixed == 1
    cplx = spm.is_complex
    dt = {(False, False): np.float64, (False, True): np.float32, (True, False): np.complex128,
          (True, True): np.complex64}[(cplx, single)]
    eps = np.finfo(np.float32 if single else np.float64).eps
    s = px.solver_export(pd)
    bc = px.bcsc_export(pd, complex_=cplx)
    tau = static_pivot_threshold(dparm, dt)
    if facto == px.PastixFactGETRF:
        F = orc.OracleLU(s, bc, tau, dt)
        G = copy.copy(F)
        G.L = px.get_coeftab(pd, s["coefnbr"], dt, side=0)
        G.U = px.get_coeftab(pd, s["coefnbr"], dt, side=1)
    else:
        F = (orc.OracleLDLT if facto == px.PastixFactLDLT else orc.OracleFactor)(s, bc, tau, dt)
        G = copy.copy(F)
        G.L = px.get_coeftab(pd, s["coefnbr"], dt)
    xo = unpermute(pd, G.solve(permute_rhs(pd, b)))
    scale = np.abs(xo).max()
    assert np.abs(x - xo).max() <= 200 * eps * scale
    assert spm.backward_error(b, x)[0] <= 10 * max(spm.backward_error(b, xo)[0], eps)
    px.finalize(pd)


def test_solve_on_a_refactorized_instance():
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


@pytest.mark.parametrize("kind,facto,mixed,nx", [("lap7", None, 0, 10), ("lap7", None, 1, 10),
                                                 ("helm", px.PastixFactLDLT, 0, 8),
                                                 ("convdiff27", px.PastixFactGETRF, 1, 8)])
def test_graph_solve_replays_match_eager(kind, facto, mixed, nx):
    """Captured passes (8 and then 3 columns: two graphs) replayed five
    times; between replays the same sweeps run eagerly (solve_graph = 0),
    and a refactorization in the middle rebuilds the task lists, after which
    the graphs are captured again.  Every replay agrees with the eager solve
    to rounding (atomic accumulation) and with it meets the backward error of
    the factor precision."""
    spm = _problem(kind, nx)
    pd, iparm, dparm = analyzed(spm, nx, facto=facto, mixed=mixed)
    px.task_numfact(pd, spm)
    eps = np.finfo(np.float32 if mixed else np.float64).eps
    for rep in range(5):
        if rep == 2:
            px.task_numfact(pd, spm)
        _, b = spm.gen_rhs(11, seed=rep)
        out = {}
        for g in (1, 0):
            px.set_option(pd, "solve_graph", g)
            x = np.array(b, order="F", copy=True)
            px.task_solve(pd, spm, x)
            out[g] = x
        scale = np.abs(out[0]).max()
        assert np.abs(out[1] - out[0]).max() <= 200 * eps * scale
        assert spm.backward_error(b, out[1])[0] <= 100 * eps
    px.finalize(pd)
