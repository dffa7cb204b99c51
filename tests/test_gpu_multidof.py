"""Multiple unknowns per node (configuration 5: 3 x 3 blocks per node of a
7-point grid, lognormal cell coefficients) on the MI355X engine: the node
ordering expanded to the unknowns (pastixOrderExpand), every tile of the
factor against the oracle, and the FP32 factor + FP64 GMRES to a backward
error of 1e-12 with many right-hand sides."""
import numpy as np
import pytest

import pastix_amd as px
import oracle_ref as orc
from helpers import analyzed, cblk_diff, make_problem, static_pivot_threshold

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("nx", [6, 9])
def test_multidof_tiles_match_oracle(nx, dtype):
    spm = make_problem("vlap7", nx)
    pd, iparm, dparm = analyzed(spm, nx, mixed=int(dtype == np.float32), dof=3)
    px.task_numfact(pd, spm)
    s = px.solver_export(pd)
    assert s["nodenbr"] == 3 * nx ** 3
    L = px.get_coeftab(pd, s["coefnbr"], dtype)
    F = orc.OracleFactor(s, px.bcsc_export(pd), static_pivot_threshold(dparm, dtype), dtype)
    assert cblk_diff(s, L, F.L, dtype) <= 10.0
    px.finalize(pd)


def test_multidof_fp32_gmres_many_rhs():
    nx = 10
    spm = make_problem("vlap7", nx)
    pd, iparm, dparm = analyzed(spm, nx, mixed=1, dof=3)
    px.task_numfact(pd, spm)
    x0, b = spm.gen_rhs(32, seed=0)
    x = np.array(b, order="F", copy=True)
    px.task_solve(pd, spm, x)
    px.set_option(pd, "berr_target", 1e-12)
    px.task_refine(pd, spm, b, x)
    berr, _ = spm.backward_error(b, x)
    assert berr <= 1e-12
    px.finalize(pd)
