"""The intra-supernode row reordering (pastix_subtask_reordering) against an
independent Python restatement of the reference's algorithm
(tests/symbol_ref.py, from symbol/symbol_reordering.c), compared exactly on
the permutation it produces.

Cases: grid nested dissections (one unknown per node and three), the
built-in nested dissection on the reference's test matrices, a non-default
split level (which fills the upper lists and exercises the up-list
equality), and small stop criteria (which cap the distances)."""
import os

import numpy as np
import pytest

import pastix_amd as px
import symbol_ref

MAT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "matrix")


def _prepare(spm, grid=None, dof=1, split=None, stop=None):
    iparm, dparm = px.initParam()
    iparm[px.IPARM_VERBOSE] = px.PastixVerboseNot
    iparm[px.IPARM_FACTORIZATION] = px.PastixFactLU
    if split is not None:
        iparm[px.IPARM_REORDERING_SPLIT] = split
    if stop is not None:
        iparm[px.IPARM_REORDERING_STOP] = stop
    if grid is not None:
        iparm[px.IPARM_ORDERING] = px.PastixOrderPersonal
    pd = px.init(iparm, dparm)
    order = None
    if grid is not None:
        order = px.Order.grid(*grid)
        if dof > 1:
            order.expand(dof)
    px.subtask_order(pd, spm, order)
    px.subtask_symbfact(pd)
    return pd, iparm


def _check(pd, iparm):
    cb, bk = px.symbol_export(pd)
    _, peri, rang, tree = px.order_export(pd)
    want = symbol_ref.reorder(cb.tolist(), bk.tolist(), rang.tolist(), tree.tolist(), peri.tolist(),
                              int(iparm[px.IPARM_REORDERING_SPLIT]), int(iparm[px.IPARM_REORDERING_STOP]))
    px.subtask_reordering(pd)
    perm2, peri2, _, _ = px.order_export(pd)
    assert peri2.tolist() == want
    assert np.array_equal(perm2[peri2], np.arange(len(peri2)))
    px.finalize(pd)


@pytest.mark.parametrize("nx,ny,nz", [(6, 6, 6), (9, 7, 5), (11, 11, 3)])
def test_grid_reordering_matches_restatement(nx, ny, nz):
    spm = px.Spm.laplacian7(nx, ny, nz)
    _check(*_prepare(spm, grid=(nx, ny, nz)))


def test_multidof_grid_reordering_matches_restatement():
    spm = px.Spm.vlaplacian7(5, 5, 5, 3, 1.0, 0)
    _check(*_prepare(spm, grid=(5, 5, 5), dof=3))


@pytest.mark.parametrize("name,driver", [("orsirr.rua", "SpmDriverHB"), ("young4c.mtx", "SpmDriverMM"),
                                         ("small.rsa", "SpmDriverRSA")])
def test_reader_matrix_reordering_matches_restatement(name, driver):
    spm = px.Spm.read(px.ENUMS[driver], os.path.join(MAT, name))
    _check(*_prepare(spm))


@pytest.mark.parametrize("split,stop", [(3, None), (5, None), (0, 3), (2, 1), (0, 0)])
def test_split_and_stop_parameters_match_restatement(split, stop):
    nx = 8
    spm = px.Spm.laplacian7(nx, nx, nx)
    _check(*_prepare(spm, grid=(nx, nx, nx), split=split, stop=stop))
