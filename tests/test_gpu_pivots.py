"""Static pivoting on the MI355X engine against the CPU oracle: pivots below
the threshold tau = eps/2 ||A|| are replaced by tau (core_zpotf2sp,
kernels/core_zpotrfsp.c:68), by sign(re d) tau (core_zsytf2sp,
kernels/core_zsytrfsp.c:64; core_zgetf2sp, kernels/core_zgetrfsp.c:81-86)
and counted in IPARM_STATIC_PIVOTING (sopalin/pastix_task_sopalin.c:532).

Every 53rd unknown is cut off from the rest of the matrix (its row and
column set to explicit zeros, the structure unchanged) with a zero
diagonal, so its pivot is exactly 0 in any elimination order: the device
and the oracle must both replace it, count the same number, and agree on
every tile.  The unknowns land at different positions of the 32-column
blocks of the diagonal tiles, in the multi-workgroup panel, the
one-workgroup panel, the LDL^T panel and the LU panel."""
import numpy as np
import pytest

import pastix_amd as px
import oracle_ref as orc
from helpers import analyzed, cblk_diff, make_problem, static_pivot_threshold

pytestmark = pytest.mark.gpu


def _cut_unknowns(spm, every=53, first=7):
    cp, rp, v = spm.arrays()
    n = len(cp) - 1
    cut = np.zeros(n, dtype=bool)
    cut[first::every] = True
    cols = np.repeat(np.arange(n), np.diff(cp))
    v = v.copy()
    v[cut[cols] | cut[rp]] = 0.0
    return px.Spm.from_csc(n, cp, rp, v, mtxtype=spm.s.mtxtype), int(cut.sum())


CASES = [
    # (matrix, facto, mixed, engine options)
    ("lap7", None, 0, {}),
    ("lap7", None, 1, {}),
    ("lap7", None, 0, {"panel_mw": 0}),
    ("lap7sd", px.PastixFactLDLT, 0, {}),
    ("lap7sd", px.PastixFactLDLT, 1, {}),
    ("convdiff27", px.PastixFactLU, 0, {}),
    ("convdiff27", px.PastixFactLU, 1, {}),
]


@pytest.mark.parametrize("kind,facto,mixed,opts", CASES)
def test_replaced_pivots_match_oracle(kind, facto, mixed, opts):
    nx = 14
    spm, ncut = _cut_unknowns(make_problem(kind, nx))
    pd, iparm, dparm = analyzed(spm, nx, facto=facto, mixed=mixed)
    for key, val in opts.items():
        px.set_option(pd, key, val)
    px.task_numfact(pd, spm)
    dt = np.float32 if mixed else np.float64
    tau = static_pivot_threshold(dparm, dt)
    s = px.solver_export(pd)
    bc = px.bcsc_export(pd)
    if facto == px.PastixFactLU:
        L = px.get_coeftab(pd, s["coefnbr"], dt, side=0)
        U = px.get_coeftab(pd, s["coefnbr"], dt, side=1)
        F = orc.OracleLU(s, bc, tau, dt)
        assert cblk_diff(s, L, F.L, dt, full_diag=True) <= 10.0
        assert cblk_diff(s, U, F.U, dt, full_diag=True) <= 10.0
    else:
        L = px.get_coeftab(pd, s["coefnbr"], dt)
        F = (orc.OracleLDLT if facto == px.PastixFactLDLT else orc.OracleFactor)(s, bc, tau, dt)
        assert cblk_diff(s, L, F.L, dt, full_diag=facto == px.PastixFactLDLT) <= 10.0
    assert ncut > 0
    assert F.npiv == ncut
    assert iparm[px.IPARM_STATIC_PIVOTING] == ncut
    px.finalize(pd)
