"""Failure paths of the device engine, each forced from the host so that
every kernel still exits normally:

- the multi-workgroup panel's bounded waits: with a zero bound (engine
  option wait_ms = 0) the waits that are not satisfied at once give up, the
  kernels finish, and the factorization returns PASTIX_ERR_INTERNAL; with
  the bound restored the same instance factorizes again and matches the
  oracle;
- a matrix entry beyond the single-precision range with a single-precision
  factor: the fill detects it and the factorization returns
  PASTIX_ERR_BADPARAMETER (the reference's mixed path refuses such a
  matrix as well, sopalin/coeftab.c);
- IPARM_ITERMAX = 1: the refinement stops after one iteration, short of
  the target, and reports that iteration count (reference
  refinement/z_refine_gmres.c:75, :314 stop on itermax)."""
import ctypes

import numpy as np
import pytest

import pastix_amd as px
import oracle_ref as orc
from helpers import analyzed, cblk_diff, static_pivot_threshold

pytestmark = pytest.mark.gpu


def test_panel_wait_bound_gives_up_and_reports_internal_error():
    nx = 20
    spm = px.Spm.laplacian7(nx, nx, nx)
    pd, iparm, dparm = analyzed(spm, nx)
    try:
        px.set_option(pd, "wait_ms", 0)
        with pytest.raises(px.PastixError) as e:
            px.task_numfact(pd, spm)
        assert e.value.code == px.PASTIX_ERR_INTERNAL
    finally:
        px.set_option(pd, "wait_ms", 2000)
    px.task_numfact(pd, spm)
    s = px.solver_export(pd)
    L = px.get_coeftab(pd, s["coefnbr"], np.float64)
    F = orc.OracleFactor(s, px.bcsc_export(pd), static_pivot_threshold(dparm, np.float64))
    assert cblk_diff(s, L, F.L, np.float64) <= 10.0
    px.finalize(pd)


def test_single_factor_of_an_entry_beyond_float_range_is_refused():
    nx = 6
    base = px.Spm.laplacian7(nx, nx, nx)
    cp, rp, v = base.arrays()
    v = v.copy()
    v[0] = 1e39   # the first diagonal entry: above FLT_MAX
    spm = px.Spm.from_csc(base.n, cp, rp, v, mtxtype=base.s.mtxtype)
    pd, iparm, dparm = analyzed(spm, nx, mixed=1)
    with pytest.raises(px.PastixError) as e:
        px.task_numfact(pd, spm)
    assert e.value.code == px.PASTIX_ERR_BADPARAMETER
    px.finalize(pd)
    # the double factor takes it
    pd, iparm, dparm = analyzed(spm, nx, mixed=0)
    px.task_numfact(pd, spm)
    px.finalize(pd)


def test_refinement_stops_at_itermax():
    nx = 12
    spm = px.Spm.laplacian7(nx, nx, nx)
    pd, iparm, dparm = analyzed(spm, nx, mixed=1, extra={px.IPARM_ITERMAX: 1})
    px.task_numfact(pd, spm)
    x0, b = spm.gen_rhs(2, seed=0)
    x = np.array(b, order="F", copy=True)
    px.task_solve(pd, spm, x)
    berr0 = spm.backward_error(b, x)[0]
    px.task_refine(pd, spm, b, x)
    assert iparm[px.IPARM_NBITER] == 1
    berr1 = spm.backward_error(b, x)[0]
    assert berr1 < berr0          # one iteration helps ...
    assert berr1 > 1e-15          # ... but a single-precision factor needs more
    px.finalize(pd)
