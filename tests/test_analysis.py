"""Analysis restatement pinned without the reference: closed forms of the grid
nested dissection, brute-force block elimination for the symbolic
factorization, invariants of reordering / splitting / solver layout, and the
flop formulas."""
import numpy as np
import pytest

import pastix_amd as px
from helpers import analyzed


def _pattern(spm):
    cp, rp, v = spm.arrays()
    n = spm.n
    A = np.zeros((n, n), dtype=bool)
    for j in range(n):
        A[rp[cp[j]:cp[j + 1]], j] = True
    return A | A.T


def _stepwise(nx, ny, nz, upto):
    spm = px.Spm.laplacian7(nx, ny, nz)
    iparm, dparm = px.initParam()
    iparm[px.IPARM_ORDERING] = px.PastixOrderPersonal
    iparm[px.IPARM_FACTORIZATION] = px.PastixFactPOTRF
    pd = px.init(iparm, dparm)
    px.subtask_order(pd, spm, px.Order.grid(nx, ny, nz))
    if upto >= 1:
        px.subtask_symbfact(pd)
    if upto >= 2:
        px.subtask_reordering(pd)
    if upto >= 3:
        px.subtask_blend(pd)
    return spm, pd, iparm, dparm


# --------------------------------------------------------------------------
# grid nested dissection (order/order_grids.c)

def _grid_reference(nx, ny, nz):
    """Independent Python statement of the grid ND numbering (separator in the
    middle of the longest direction, x on ties, y before z; boxes smaller than
    15 vertices are leaves numbered x-outermost; separator first, then the lower
    half, then the upper half, all from the highest number down)."""
    perm = -np.ones(nx * ny * nz, dtype=np.int64)
    ranges, depths = [], []
    nxt = [nx * ny * nz - 1]

    def idx(i, j, k):
        return i + nx * j + nx * ny * k

    def rec(x0, x1, y0, y1, z0, z1, d):
        ranges.append(nxt[0])
        depths.append(d)
        sx, sy, sz = x1 - x0, y1 - y0, z1 - z0
        if sx * sy * sz < 15:
            for i in range(x0, x1):
                for j in range(y0, y1):
                    for k in range(z0, z1):
                        perm[idx(i, j, k)] = nxt[0]
                        nxt[0] -= 1
            return
        dr = 0
        if sy > sx:
            dr = 1
        if sz > sx and sz > sy:
            dr = 2
        if dr == 0:
            m = x0 + sx // 2
            for j in range(y0, y1):
                for k in range(z0, z1):
                    perm[idx(m, j, k)] = nxt[0]
                    nxt[0] -= 1
            rec(x0, m, y0, y1, z0, z1, d + 1)
            rec(m + 1, x1, y0, y1, z0, z1, d + 1)
        elif dr == 1:
            m = y0 + sy // 2
            for i in range(x0, x1):
                for k in range(z0, z1):
                    perm[idx(i, m, k)] = nxt[0]
                    nxt[0] -= 1
            rec(x0, x1, y0, m, z0, z1, d + 1)
            rec(x0, x1, m + 1, y1, z0, z1, d + 1)
        else:
            m = z0 + sz // 2
            for i in range(x0, x1):
                for j in range(y0, y1):
                    perm[idx(i, j, m)] = nxt[0]
                    nxt[0] -= 1
            rec(x0, x1, y0, y1, z0, m, d + 1)
            rec(x0, x1, y0, y1, m + 1, z1, d + 1)

    rec(0, nx, 0, ny, 0, nz, 1)
    c = len(ranges)
    rang = np.array([0] + [ranges[c - 1 - i] + 1 for i in range(c)])
    dep = [depths[c - 1 - i] for i in range(c)]
    tree = []
    for i in range(c):
        f = next((j for j in range(i + 1, c) if dep[j] < dep[i]), -1)
        tree.append(f)
    return perm, rang, np.array(tree)


@pytest.mark.parametrize("dims", [(3, 3, 3), (5, 4, 3), (7, 7, 7), (10, 6, 9), (15, 15, 15)])
def test_grid_order_matches_closed_form(dims):
    perm, peri, rang, tree = px.Order.grid(*dims).arrays()
    rp, rr, rt = _grid_reference(*dims)
    assert np.array_equal(perm, rp)
    assert np.array_equal(rang, rr)
    assert np.array_equal(tree, rt)
    assert np.array_equal(peri[perm], np.arange(len(perm)))


def test_grid_order_separators_disconnect_the_halves():
    nx = 9
    spm = px.Spm.laplacian7(nx, nx, nx)
    A = _pattern(spm)
    perm, peri, rang, tree = px.Order.grid(nx, nx, nx).arrays()
    top = np.arange(rang[-2], rang[-1])            # last supernode = top separator
    rest = np.setdiff1d(np.arange(nx ** 3), peri[top])
    sub = A[np.ix_(rest, rest)]
    # breadth-first search: removing the separator leaves two components
    seen = np.zeros(len(rest), dtype=bool)
    comps = 0
    for s in range(len(rest)):
        if seen[s]:
            continue
        comps += 1
        stack = [s]
        seen[s] = True
        while stack:
            u = stack.pop()
            for w in np.nonzero(sub[u])[0]:
                if not seen[w]:
                    seen[w] = True
                    stack.append(w)
    assert comps == 2


# --------------------------------------------------------------------------
# symbolic factorization (symbol/symbol_fax.c) by brute-force elimination

def _brute_symbol(A, perm, rang):
    """Block elimination on booleans: the rows below a supernode are the
    nonzeros of its columns after the updates of all earlier supernodes; each
    supernode then fills the dense square of its rows."""
    n = A.shape[0]
    P = np.zeros_like(A)
    P[np.ix_(perm, perm)] = A
    M = np.tril(P) | np.eye(n, dtype=bool)
    rows = []
    for k in range(len(rang) - 1):
        f, l = rang[k], rang[k + 1] - 1
        R = np.nonzero(M[l + 1:, f:l + 1].any(axis=1))[0] + l + 1
        rows.append(R)
        if len(R):
            M[np.ix_(R, R)] |= np.tril(np.ones((len(R), len(R)), dtype=bool))
    return rows


@pytest.mark.parametrize("dims,reorder", [((5, 5, 5), False), ((6, 4, 5), True), ((8, 8, 8), True)])
def test_symbol_equals_block_elimination(dims, reorder):
    spm, pd, iparm, dparm = _stepwise(*dims, upto=2 if reorder else 1)
    perm, peri, rang, tree = px.order_export(pd)
    cb, bk = px.symbol_export(pd)
    assert len(cb) - 1 == len(rang) - 1
    brute = _brute_symbol(_pattern(spm), perm, rang)
    for k in range(len(rang) - 1):
        assert cb[k, 0] == rang[k] and cb[k, 1] == rang[k + 1] - 1
        bl = bk[cb[k, 2]:cb[k + 1, 2]]
        assert bl[0, 0] == rang[k] and bl[0, 1] == rang[k + 1] - 1 and bl[0, 3] == k
        got = np.concatenate([np.arange(r[0], r[1] + 1) for r in bl[1:]]) if len(bl) > 1 else np.zeros(0)
        assert np.array_equal(got, brute[k])
        # bloks are maximal runs inside one facing cblk
        for r in bl[1:]:
            assert rang[r[3]] <= r[0] and r[1] < rang[r[3] + 1]
        for a, b in zip(bl[1:-1], bl[2:]):
            assert not (b[0] == a[1] + 1 and b[3] == a[3])
    px.finalize(pd)


def test_reordering_keeps_supernodes_and_only_permutes_inside_them():
    spm, pd, _, _ = _stepwise(9, 9, 9, upto=1)
    perm0, peri0, rang0, _ = px.order_export(pd)
    nb0 = px.symbol_export(pd)[1].shape[0]
    px.subtask_reordering(pd)
    perm1, peri1, rang1, _ = px.order_export(pd)
    nb1 = px.symbol_export(pd)[1].shape[0]
    assert np.array_equal(rang0, rang1)
    for k in range(len(rang0) - 1):
        a, b = rang0[k], rang0[k + 1]
        assert set(peri0[a:b]) == set(peri1[a:b])
    assert not np.array_equal(peri0, peri1)
    assert nb1 <= nb0          # the TSP ordering compacts the off-diagonal blocks
    px.finalize(pd)


# --------------------------------------------------------------------------
# splitting and solver layout (blend/splitsymbol.c, blend/solver_matrix_gen*.c)

def test_split_widths_and_solver_layout():
    nx = 24
    spm, pd, iparm, dparm = _stepwise(nx, nx, nx, upto=2)
    cb0, _ = px.symbol_export(pd)
    px.subtask_blend(pd)
    s = px.solver_export(pd)
    cb = s["cblk"].reshape(-1, 6)
    bk = s["blok"].reshape(-1, 5)
    w = cb[:-1, 1] - cb[:-1, 0] + 1
    # wide supernodes are cut into pieces of at least min_blocksize columns
    w0 = cb0[:-1, 1] - cb0[:-1, 0] + 1
    assert w0.max() > iparm[px.IPARM_MAX_BLOCKSIZE]
    for k in range(len(w)):
        if w[k] > iparm[px.IPARM_MAX_BLOCKSIZE]:
            pytest.fail("cblk %d wider than max" % k)
    # every original supernode is a union of consecutive pieces
    starts = set(cb[:-1, 0].tolist())
    assert set(cb0[:-1, 0].tolist()) <= starts
    # tile layout: coefind = rows above * ncols, stride = all rows, coefoff cumulative
    coef = 0
    for k in range(s["cblknbr"]):
        n = cb[k, 1] - cb[k, 0] + 1
        rows = 0
        assert cb[k, 4] == coef
        for b in range(cb[k, 2], cb[k + 1, 2]):
            assert bk[b, 4] == rows * n and bk[b, 2] == k
            rows += bk[b, 1] - bk[b, 0] + 1
        assert rows == cb[k, 3]
        coef += rows * n
    assert coef == s["coefnbr"]
    px.finalize(pd)


# --------------------------------------------------------------------------
# flops (symbol/symbol_cost.c with common/flops.h)

def _flops_cholesky(cb, bk):
    tot = 0.0
    for k in range(len(cb) - 1):
        N = float(cb[k, 1] - cb[k, 0] + 1)
        M = float(sum(bk[b, 1] - bk[b, 0] + 1 for b in range(cb[k, 2] + 1, cb[k + 1, 2])))
        potrf = N * (((1 / 6) * N + 0.5) * N + 1 / 3) + N * (((1 / 6) * N) * N - 1 / 6)
        trsm = 2 * 0.5 * M * N * (N + 1)          # adds counted as multiplies
        syrk = 2 * 0.5 * N * M * (M + 1)
        tot += potrf + (trsm + syrk if M > 0 else 0)
    return tot


def test_thflops_closed_form():
    spm, pd, iparm, dparm = _stepwise(12, 12, 12, upto=3)
    s = px.solver_export(pd)
    cb = s["cblk"].reshape(-1, 6)[:, [0, 1, 2]]
    bk = s["blok"].reshape(-1, 5)
    assert dparm[px.DPARM_FACT_THFLOPS] == pytest.approx(_flops_cholesky(cb, bk), rel=1e-12)
    # a dense 1-cblk matrix: THFLOPS = n^3/3 + n^2/2 + n/6
    spm1 = px.Spm.laplacian7(4, 1, 1)
    iparm1, dparm1 = px.initParam()
    iparm1[px.IPARM_ORDERING] = px.PastixOrderPersonal
    iparm1[px.IPARM_FACTORIZATION] = px.PastixFactPOTRF
    pd1 = px.init(iparm1, dparm1)
    px.analyze_grid(pd1, spm1, 4, 1, 1)
    n = 4.0
    assert dparm1[px.DPARM_FACT_THFLOPS] == pytest.approx(n ** 3 / 3 + n ** 2 / 2 + n / 6)
    px.finalize(pd)
    px.finalize(pd1)


def test_level_order_puts_children_before_fathers():
    spm, pd, _, _ = _stepwise(16, 16, 16, upto=0)
    _, _, rang, tree = px.order_export(pd)
    assert tree[-1] == -1
    for k, f in enumerate(tree):
        assert f == -1 or f > k
    px.finalize(pd)


def test_order_expand_for_multiple_dof():
    """pastixOrderExpand (reference order/order.c:396-469): every node's dof
    unknowns take consecutive positions in the node's place, ranges scale by
    dof, the tree is kept; and the analysis of the expanded matrix is the
    node analysis with every column and row range scaled by dof (what the
    reference's symbol expansion gives) while no cblk needs a split."""
    from helpers import make_problem, analyzed
    o = px.Order.grid(6, 5, 4)
    p0, i0, r0, t0 = o.arrays()
    o.expand(3)
    p, i, r, t = o.arrays()
    assert np.array_equal(i, (3 * i0[:, None] + np.arange(3)).ravel())
    assert np.array_equal(p[i], np.arange(len(i)))
    assert np.array_equal(r, 3 * r0) and np.array_equal(t, t0)
    assert px.lib.pastixOrderCheck(o.ptr) == 0
    for nx in (6, 8):
        a = make_problem("lap7", nx)
        pa, _, _ = analyzed(a, nx)
        b = make_problem("vlap7", nx)
        pb, _, _ = analyzed(b, nx, dof=3)
        sa, sb = px.solver_export(pa), px.solver_export(pb)
        assert sb["nodenbr"] == 3 * sa["nodenbr"] and sb["cblknbr"] == sa["cblknbr"]
        ca, cb = sa["cblk"].reshape(-1, 6), sb["cblk"].reshape(-1, 6)
        assert np.array_equal(cb[:-1, 0], 3 * ca[:-1, 0]) and np.array_equal(cb[:-1, 1] + 1, 3 * (ca[:-1, 1] + 1))
        ba, bb = sa["blok"].reshape(-1, 5), sb["blok"].reshape(-1, 5)
        assert np.array_equal(bb[:, 0], 3 * ba[:, 0]) and np.array_equal(bb[:, 1] + 1, 3 * (ba[:, 1] + 1))
        assert np.array_equal(bb[:, 2:4], ba[:, 2:4])
        px.finalize(pa)
        px.finalize(pb)
