"""The CPU oracle pinned against what the mathematics fixes: dense LAPACK-style
Cholesky of P A P^T (tile by tile, zero outside the symbol), hand-computed
static-pivot cases, exact SpMV on integer-valued matrices, solve and GMRES
convergence, and the plrnt generator against sequential LCG stepping."""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp

import pastix_amd as px
import oracle_ref as orc
from helpers import analyzed, make_problem, permute_rhs, static_pivot_threshold


def _dense(cp, rt, v, n):
    A = np.zeros((n, n), dtype=v.dtype)
    for j in range(n):
        A[rt[cp[j]:cp[j + 1]], j] = v[cp[j]:cp[j + 1]]
    return A


def _setup(nx, ny=None, nz=None):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    spm = px.Spm.laplacian7(nx, ny, nz)
    pd, iparm, dparm = analyzed(spm, nx, ny, nz)
    px._check("spm2bcsc", px.lib.pastix_subtask_spm2bcsc(pd, ctypes.byref(spm.s)))
    return spm, pd, iparm, dparm, px.solver_export(pd), px.bcsc_export(pd)


@pytest.mark.parametrize("dims", [(4, 4, 4), (7, 5, 6), (10, 10, 10)])
@pytest.mark.parametrize("nthreads", [1, 4])
def test_oracle_cholesky_equals_dense(dims, nthreads):
    spm, pd, iparm, dparm, s, bc = _setup(*dims)
    n = s["nodenbr"]
    F = orc.OracleFactor(s, bc, static_pivot_threshold(dparm, np.float64), nthreads=nthreads)
    A = _dense(*bc, n)
    Ld = np.linalg.cholesky(A)
    Lo = orc.dense_from_factor(s, F.L)
    assert np.abs(Lo - Ld).max() <= 1e-13
    # the symbolic structure holds every nonzero of the factor
    assert np.abs(Ld[Lo == 0]).max() <= 1e-13
    assert np.abs(Lo @ Lo.T - A).max() <= 1e-12 * np.abs(A).max()
    px.finalize(pd)


def test_oracle_single_precision_factor():
    spm, pd, iparm, dparm, s, bc = _setup(6)
    n = s["nodenbr"]
    F = orc.OracleFactor(s, bc, static_pivot_threshold(dparm, np.float32), np.float32)
    Ld = np.linalg.cholesky(_dense(*bc, n))
    Lo = orc.dense_from_factor(s, F.L).astype(np.float64)
    assert np.abs(Lo - Ld).max() <= 50 * np.finfo(np.float32).eps * np.abs(Ld).max()
    px.finalize(pd)


def _one_cblk(n):
    """Solver arrays of a single dense cblk of width n."""
    cblk = np.array([0, n - 1, 0, n, 0, 0, n, n, 1, 0, n * n, 0], dtype=np.int64)
    blok = np.array([0, n - 1, 0, 0, 0], dtype=np.int64)
    return dict(cblknbr=1, bloknbr=1, coefnbr=n * n, nodenbr=n, cblk=cblk, blok=blok,
                browtab=np.zeros(1, dtype=np.int64))


def _csc(A):
    M = sp.csc_matrix(A)
    M.sort_indices()
    return (M.indptr.astype(np.int64), M.indices.astype(np.int64), M.data.astype(np.float64))


def test_static_pivot_known_answers():
    # |a_00| < tau is replaced by tau (core_zpotf2sp), then L = sqrt
    tau = 1e-8
    A = np.array([[1e-20, 0., 0.], [0., 4., 2.], [0., 2., 5.]])
    F = orc.OracleFactor(_one_cblk(3), _csc(A), tau)
    L = F.L.reshape(3, 3).T
    assert F.npiv == 1
    assert L[0, 0] == pytest.approx(np.sqrt(tau))
    assert L[1, 1] == pytest.approx(2.0) and L[2, 1] == pytest.approx(1.0) and L[2, 2] == pytest.approx(2.0)
    # a pivot exactly at the threshold is kept
    A2 = np.diag([tau, 9.0])
    F2 = orc.OracleFactor(_one_cblk(2), _csc(A2), tau)
    assert F2.npiv == 0 and F2.L[0] == pytest.approx(np.sqrt(tau))
    # a pivot that becomes tiny during elimination is counted too
    A3 = np.array([[1.0, 1.0], [1.0, 1.0 + 1e-18]])
    F3 = orc.OracleFactor(_one_cblk(2), _csc(A3), tau)
    assert F3.npiv == 1 and F3.L.reshape(2, 2).T[1, 1] == pytest.approx(np.sqrt(tau))


def test_oracle_spmv_is_exact_on_the_integer_laplacian():
    spm, pd, iparm, dparm, s, bc = _setup(8)
    cp, rt, v = bc
    x = np.round(px.plrnt(s["nodenbr"], seed=7) * 64)      # small integers: exact arithmetic
    y = orc.spmv(cp, rt, v, x)
    M = sp.csc_matrix((v, rt, cp), shape=(s["nodenbr"],) * 2)
    assert np.array_equal(y, M.T @ x)
    assert np.array_equal(y, M @ x)                         # symmetric storage is full
    px.finalize(pd)


@pytest.mark.parametrize("complex_", [False, True])
def test_backward_error_reports_nan(complex_):
    """A NaN anywhere in x makes its column's backward error NaN (a max that
    drops NaN would report the column as solved)."""
    spm = px.Spm.laplacian7(4, 4, 4, complex_=complex_)
    x0, b = spm.gen_rhs(2, seed=0)
    worst, each = spm.backward_error(b, x0)
    assert worst <= 1e-15
    x = np.array(x0, copy=True)
    x[5, 1] = np.nan
    worst, each = spm.backward_error(b, x)
    assert np.isnan(worst) and each[0] <= 1e-15 and np.isnan(each[1])


@pytest.mark.parametrize("nrhs", [1, 3])
def test_oracle_solve_backward_error(nrhs):
    spm, pd, iparm, dparm, s, bc = _setup(12)
    F = orc.OracleFactor(s, bc, static_pivot_threshold(dparm, np.float64))
    x0, b = spm.gen_rhs(nrhs, seed=0)
    xp = F.solve(permute_rhs(pd, b))
    perm = px.order_export(pd)[0]
    x = xp[perm]
    berr, _ = spm.backward_error(b, x)
    assert berr <= 1e-14
    assert np.abs(x - x0).max() <= 1e-12
    px.finalize(pd)


def test_oracle_gmres_with_single_factor_converges():
    spm, pd, iparm, dparm, s, bc = _setup(10)
    F32 = orc.OracleFactor(s, bc, static_pivot_threshold(dparm, np.float32), np.float32)
    x0, b = spm.gen_rhs(1, seed=2)
    bp = permute_rhs(pd, b)[:, 0]
    xs = F32.solve(bp)
    r0 = np.linalg.norm(bp - orc.spmv(*bc, xs)) / np.linalg.norm(bp)
    assert r0 > 1e-12
    x, it = F32.gmres(bp, xs)
    assert np.linalg.norm(bp - orc.spmv(*bc, x)) / np.linalg.norm(bp) <= 1e-12
    assert 1 <= it <= 10
    # with an exact (double) preconditioner the first residual already passes
    F64 = orc.OracleFactor(s, bc, static_pivot_threshold(dparm, np.float64))
    x64, it64 = F64.gmres(bp, F64.solve(bp))
    assert it64 == 0
    px.finalize(pd)


def test_plrnt_jump_matches_sequential_lcg():
    a, c, m = 6364136223846793005, 1, 1 << 64
    seed = 12345
    ran, seq = seed, []
    for _ in range(50):
        seq.append(np.float32(0.5) - np.float32(ran) * np.float32(5.4210108624275222e-20))
        ran = (a * ran + c) % m
    assert np.array_equal(px.plrnt(50, seed=seed), np.array(seq, dtype=np.float64))
    # the O(log n) jump lands on the same stream
    assert np.array_equal(px.plrnt(20, seed=seed, shift=30), np.array(seq[30:50], dtype=np.float64))
    v = px.plrnt(100000, seed=0)
    assert v.min() >= -0.5 and v.max() <= 0.5 and abs(v.mean()) < 0.01


def _dense_lu_nopiv(A):
    """Doolittle LU without pivoting, packed (unit L strictly below, U on and above)."""
    A = np.array(A, copy=True)
    n = A.shape[0]
    for k in range(n):
        A[k + 1:, k] /= A[k, k]
        A[k + 1:, k + 1:] -= np.outer(A[k + 1:, k], A[k, k + 1:])
    return A


@pytest.mark.parametrize("dims", [(4, 4, 4), (6, 5, 4), (7, 7, 7)])
def test_oracle_lu_equals_dense(dims):
    spm = px.Spm.convdiff27(*dims, 1.0, 0.5, 3)
    pd, iparm, dparm = analyzed(spm, *dims, facto=px.PastixFactLU)
    px._check("spm2bcsc", px.lib.pastix_subtask_spm2bcsc(pd, ctypes.byref(spm.s)))
    s, bc = px.solver_export(pd), px.bcsc_export(pd)
    n = s["nodenbr"]
    A = _dense(*bc, n)
    assert np.abs(A - A.T).max() > 0.01          # the values are not symmetric
    F = orc.OracleLU(s, bc, static_pivot_threshold(dparm, np.float64))
    assert F.npiv == 0
    D = _dense_lu_nopiv(A)
    Lo = orc.dense_from_factor(s, F.L)
    Uo = orc.dense_from_factor(s, F.U)
    packed = np.tril(Lo, -1) + np.triu(Uo.T)
    assert np.abs(packed - D).max() <= 1e-12 * np.abs(D).max()
    # the symbol holds every nonzero of the factors
    assert np.abs(D[(np.tril(Lo, -1) == 0) & (np.tril(np.ones_like(D), -1) > 0)]).max() <= 1e-13
    Ld = np.tril(D, -1) + np.eye(n)
    assert np.abs(Ld @ np.triu(D) - A).max() <= 1e-12 * np.abs(A).max()
    # solve
    x0, b = spm.gen_rhs(2, seed=1)
    xp = F.solve(permute_rhs(pd, b))
    x = xp[px.order_export(pd)[0]]
    assert np.abs(x - x0).max() <= 1e-11
    px.finalize(pd)


def test_lu_static_pivot_known_answers():
    # |a_00| < tau is replaced by sign(a_00) tau (core_zgetf2sp), no square root
    tau = 1e-8
    A = np.array([[-1e-20, 1.0, 0.0], [2.0, 3.0, 1.0], [0.0, 1.0, 4.0]])
    F = orc.OracleLU(_one_cblk(3), _csc(A), tau)
    assert F.npiv == 1
    P = F.L.reshape(3, 3).T                    # packed LU in the L diagonal tile
    assert P[0, 0] == -tau
    assert P[1, 0] == pytest.approx(2.0 / -tau)
    assert P[1, 1] == pytest.approx(3.0 - (2.0 / -tau) * 1.0)
    # the U tile holds the packed factor transposed
    assert np.array_equal(F.U.reshape(3, 3), P)
    # a pivot that is exactly tau is kept, a positive tiny one becomes +tau
    F2 = orc.OracleLU(_one_cblk(2), _csc(np.array([[1e-30, 0.0], [0.0, 1.0]])), tau)
    assert F2.npiv == 1 and F2.L[0] == tau


def _dense_ldlt_nopiv(A):
    """LDL^T without pivoting: returns (unit lower L, d)."""
    A = np.array(A, copy=True)
    n = A.shape[0]
    Lm = np.eye(n, dtype=A.dtype)
    d = np.zeros(n, dtype=A.dtype)
    for k in range(n):
        d[k] = A[k, k]
        Lm[k + 1:, k] = A[k + 1:, k] / d[k]
        A[k + 1:, k + 1:] -= d[k] * np.outer(Lm[k + 1:, k], Lm[k + 1:, k])
    return Lm, d


@pytest.mark.parametrize("dims", [(4, 4, 4), (6, 5, 4), (8, 8, 8)])
def test_oracle_ldlt_equals_dense(dims):
    # shifted Laplacian: symmetric indefinite (the smallest eigenvalue of the
    # 7-point stencil is below 1.7 on these grids) and far from singular
    spm = px.Spm.laplacian7(*dims, 4.3, 1.0)
    pd, iparm, dparm = analyzed(spm, *dims, facto=px.PastixFactLDLT)
    px._check("spm2bcsc", px.lib.pastix_subtask_spm2bcsc(pd, ctypes.byref(spm.s)))
    s, bc = px.solver_export(pd), px.bcsc_export(pd)
    n = s["nodenbr"]
    A = _dense(*bc, n)
    ev = np.linalg.eigvalsh(A)
    assert ev.min() < 0 < ev.max() and np.abs(ev).min() > 0.1
    F = orc.OracleLDLT(s, bc, static_pivot_threshold(dparm, np.float64))
    assert F.npiv == 0
    Ld, d = _dense_ldlt_nopiv(A)
    Lo = orc.dense_from_factor(s, F.L)
    assert np.abs(np.tril(Lo, -1) - np.tril(Ld, -1)).max() <= 1e-11
    assert np.abs(np.diag(Lo) - d).max() <= 1e-12 * np.abs(d).max()
    # the upper part of the diagonal tiles holds D L^T
    cb = s["cblk"].reshape(-1, 6)
    for k in range(s["cblknbr"]):
        nk = cb[k, 1] - cb[k, 0] + 1
        T = F.L[cb[k, 4]:cb[k, 4] + nk * nk].reshape(nk, nk).T
        assert np.allclose(np.triu(T, 1), np.triu((np.tril(T, -1) * np.diag(T)[None, :]).T, 1), atol=1e-11)
    x0, b = spm.gen_rhs(2, seed=2)
    xp = F.solve(permute_rhs(pd, b))
    x = xp[px.order_export(pd)[0]]
    assert np.abs(x - x0).max() <= 1e-9 * np.abs(x0).max() * np.linalg.cond(A)
    berr, _ = spm.backward_error(b, x)
    assert berr <= 1e-13
    px.finalize(pd)


def test_ldlt_static_pivot_known_answers():
    # |a_00| < tau -> sign(a_00) tau, no square root; D L^T above the diagonal
    tau = 1e-8
    A = np.array([[-1e-20, 2.0, 0.0], [2.0, 3.0, 1.0], [0.0, 1.0, 4.0]])
    F = orc.OracleLDLT(_one_cblk(3), _csc(A), tau)
    T = F.L.reshape(3, 3).T
    assert F.npiv == 1
    assert T[0, 0] == -tau
    assert T[1, 0] == pytest.approx(2.0 / -tau)
    assert T[0, 1] == pytest.approx(2.0)                       # (D L^T)[0,1] = d0 * l10
    assert T[1, 1] == pytest.approx(3.0 - (2.0 / -tau) ** 2 * -tau)


@pytest.mark.parametrize("dims", [(4, 4, 4), (6, 5, 4), (8, 7, 6)])
def test_oracle_complex_symmetric_ldlt_equals_dense(dims):
    # A = L7 - (0.5 + 0.1i) I: complex symmetric, not Hermitian; the
    # factorization uses plain transposes (core_zsytrfsp)
    spm = px.Spm.helmholtz7(*dims, 0.5, 0.1)
    pd, iparm, dparm = analyzed(spm, *dims, facto=px.PastixFactLDLT)
    px._check("spm2bcsc", px.lib.pastix_subtask_spm2bcsc(pd, ctypes.byref(spm.s)))
    s, bc = px.solver_export(pd), px.bcsc_export(pd, complex_=True)
    n = s["nodenbr"]
    A = _dense(*bc, n)
    assert np.abs(A - A.T).max() == 0 and np.abs(A - A.conj().T).max() > 0.1
    F = orc.OracleLDLT(s, bc, static_pivot_threshold(dparm, np.float64), np.complex128)
    assert F.npiv == 0
    Ld, d = _dense_ldlt_nopiv(A)
    Lo = orc.dense_from_factor(s, F.L)
    assert np.abs(np.tril(Lo, -1) - np.tril(Ld, -1)).max() <= 1e-11
    assert np.abs(np.diag(Lo) - d).max() <= 1e-12 * np.abs(d).max()
    assert np.abs((np.tril(Ld, -1) + np.eye(n)) @ np.diag(d) @ (np.tril(Ld, -1) + np.eye(n)).T - A).max() <= 1e-11
    x0, b = spm.gen_rhs(2, seed=3)
    xp = F.solve(permute_rhs(pd, b))
    x = xp[px.order_export(pd)[0]]
    berr, _ = spm.backward_error(b, x)
    assert berr <= 1e-13
    # the single precision factor
    F32 = orc.OracleLDLT(s, bc, static_pivot_threshold(dparm, np.float32), np.complex64)
    L32 = orc.dense_from_factor(s, F32.L).astype(np.complex128)
    assert np.abs(np.tril(L32, -1) - np.tril(Ld, -1)).max() <= 1e-4
    px.finalize(pd)


def test_oracle_complex_lu_equals_dense():
    dims = (5, 4, 4)
    spm = px.Spm.helmholtz7(*dims, 0.5, 0.1)
    pd, iparm, dparm = analyzed(spm, *dims, facto=px.PastixFactLU)
    px._check("spm2bcsc", px.lib.pastix_subtask_spm2bcsc(pd, ctypes.byref(spm.s)))
    s, bc = px.solver_export(pd), px.bcsc_export(pd, complex_=True)
    A = _dense(*bc, s["nodenbr"])
    F = orc.OracleLU(s, bc, static_pivot_threshold(dparm, np.float64), np.complex128)
    D = _dense_lu_nopiv(A.astype(np.complex128))
    packed = np.tril(orc.dense_from_factor(s, F.L), -1) + np.triu(orc.dense_from_factor(s, F.U).T)
    assert np.abs(packed - D).max() <= 1e-12 * np.abs(D).max()
    px.finalize(pd)


@pytest.mark.parametrize("dims", [(4, 4, 4), (6, 5, 4)])
def test_oracle_single_lu_equals_dense(dims):
    """The single-precision LU instantiation against the dense no-pivot LU
    of P A P^T in double: within the growth-free bound of a few hundred
    eps_32 (a convection-diffusion matrix is diagonally dominant)."""
    spm = px.Spm.convdiff27(*dims, 1.0, 0.5, 3)
    pd, iparm, dparm = analyzed(spm, *dims, facto=px.PastixFactLU)
    px._check("spm2bcsc", px.lib.pastix_subtask_spm2bcsc(pd, ctypes.byref(spm.s)))
    s, bc = px.solver_export(pd), px.bcsc_export(pd)
    n = s["nodenbr"]
    A = _dense(*bc, n)
    F = orc.OracleLU(s, bc, static_pivot_threshold(dparm, np.float32), np.float32)
    assert F.npiv == 0 and F.L.dtype == np.float32
    D = _dense_lu_nopiv(A)
    Lo = orc.dense_from_factor(s, F.L.astype(np.float64))
    Uo = orc.dense_from_factor(s, F.U.astype(np.float64))
    packed = np.tril(Lo, -1) + np.triu(Uo.T)
    eps = np.finfo(np.float32).eps
    assert np.abs(packed - D).max() <= 300 * eps * np.abs(D).max()
    # a wrong sign or a transposed operand would be off by O(1), not eps
    assert np.abs(packed - D).max() > 0
    x0, b = spm.gen_rhs(2, seed=1)
    x = F.solve(permute_rhs(pd, b))[px.order_export(pd)[0]]
    assert np.abs(x - x0).max() <= 1e3 * eps * np.abs(x0).max()
    px.finalize(pd)


@pytest.mark.parametrize("dims", [(4, 4, 4), (6, 5, 4)])
def test_oracle_single_ldlt_equals_dense(dims):
    """The real single-precision LDL^T instantiation against the dense LDL^T
    in double of a symmetric indefinite, diagonally dominant stencil
    (diagonal +7 / -7 on alternate vertices): its Schur complements stay
    diagonally dominant, so the error stays at the level of eps_32."""
    spm = make_problem("lap7sd", *dims)
    pd, iparm, dparm = analyzed(spm, *dims, facto=px.PastixFactLDLT)
    px._check("spm2bcsc", px.lib.pastix_subtask_spm2bcsc(pd, ctypes.byref(spm.s)))
    s, bc = px.solver_export(pd), px.bcsc_export(pd)
    n = s["nodenbr"]
    A = _dense(*bc, n)
    F = orc.OracleLDLT(s, bc, static_pivot_threshold(dparm, np.float32), np.float32)
    assert F.npiv == 0 and F.L.dtype == np.float32
    Ld, d = _dense_ldlt_nopiv(A)
    Lo = orc.dense_from_factor(s, F.L.astype(np.float64))
    # without pivoting an indefinite matrix amplifies rounding (small d's):
    # the double factor measures the amplification, the single one must stay
    # within it, scaled by the ratio of the epsilons
    F64 = orc.OracleLDLT(s, bc, static_pivot_threshold(dparm, np.float64), np.float64)
    L64 = orc.dense_from_factor(s, F64.L)
    eps, eps64 = np.finfo(np.float32).eps, np.finfo(np.float64).eps
    amp = max(np.abs(np.tril(L64, -1) - np.tril(Ld, -1)).max() / eps64, 1.0)
    err = np.abs(np.tril(Lo, -1) - np.tril(Ld, -1)).max()
    assert 0 < err <= 20 * amp * eps
    ampd = max(np.abs(np.diag(L64) - d).max() / (eps64 * np.abs(d).max()), 1.0)
    assert np.abs(np.diag(Lo) - d).max() <= 20 * ampd * eps * np.abs(d).max()
    x0, b = spm.gen_rhs(2, seed=2)
    x = F.solve(permute_rhs(pd, b))[px.order_export(pd)[0]]
    berr, _ = spm.backward_error(b, x)
    assert berr <= 20 * amp * eps
    px.finalize(pd)
