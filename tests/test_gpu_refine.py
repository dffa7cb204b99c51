"""Iterative refinement variants on the device (GPU tests): CG for the SPD
Laplacian (reference refinement/z_refine_grad.c) and BiCGStab for the
unsymmetric convection-diffusion operator (refinement/z_refine_bicgstab.c),
both preconditioned by a single-precision factor, to a backward error of
1e-12; with an exact (double) factor both stop after one iteration."""
import numpy as np
import pytest

import pastix_amd as px
from helpers import analyzed, make_problem

pytestmark = pytest.mark.gpu


def _run(kind, nx, facto, mixed, refine, nrhs=2, seed=0):
    spm = make_problem(kind, nx, seed=1) if kind == "convdiff27" else make_problem(kind, nx)
    pd, iparm, dparm = analyzed(spm, nx, facto=facto, mixed=mixed,
                                extra={px.IPARM_REFINEMENT: refine})
    px.task_numfact(pd, spm)
    x0, b = spm.gen_rhs(nrhs, seed=seed)
    x = np.array(b, order="F", copy=True)
    px.task_solve(pd, spm, x)
    px.set_option(pd, "berr_target", 1e-12)
    px.task_refine(pd, spm, b, x)
    berr, _ = spm.backward_error(b, x)
    it = int(iparm[px.IPARM_NBITER])
    px.finalize(pd)
    return berr, it


def test_cg_single_factor_laplacian():
    berr, it = _run("lap7", 14, px.PastixFactPOTRF, 1, px.PastixRefineCG, nrhs=3)
    assert berr <= 1e-12 and 0 < it < 30


def test_bicgstab_single_factor_lu():
    berr, it = _run("convdiff27", 10, px.PastixFactLU, 1, px.PastixRefineBiCGSTAB, nrhs=2)
    assert berr <= 1e-12 and 0 < it < 30


def test_bicgstab_single_factor_cholesky():
    berr, it = _run("lap7", 12, px.PastixFactPOTRF, 1, px.PastixRefineBiCGSTAB, nrhs=1, seed=5)
    assert berr <= 1e-12 and 0 < it < 30


@pytest.mark.parametrize("refine", ["CG", "BiCGSTAB", "SR"])
def test_exact_factor_needs_at_most_one_iteration(refine):
    # the double factor's solve is already at 1e-12 after one correction
    kind, facto = ("convdiff27", px.PastixFactLU) if refine == "BiCGSTAB" else ("lap7", px.PastixFactPOTRF)
    berr, it = _run(kind, 9, facto, 0, getattr(px, "PastixRefine" + refine))
    assert berr <= 1e-12 and it <= 1
