"""Shared helpers of the test-suite: problem set-up through the PaStiX API."""
import numpy as np

import pastix_amd as px


def make_problem(kind, nx, ny=None, nz=None, **kw):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    if kind == "lap7":
        return px.Spm.laplacian7(nx, ny, nz, kw.get("diag", 6.0), kw.get("offd", 1.0))
    if kind == "convdiff27":
        return px.Spm.convdiff27(nx, ny, nz, 1.0, 0.5, kw.get("seed", 0))
    if kind == "helmholtz7":
        return px.Spm.helmholtz7(nx, ny, nz, 0.5, 0.1)
    if kind == "vlap7":
        # configuration 5: dof x dof blocks per node, lognormal coefficients
        return px.Spm.vlaplacian7(nx, ny, nz, kw.get("dof", 3), kw.get("sigma", 1.0), kw.get("seed", 0))
    if kind == "lap7sd":
        # symmetric indefinite and diagonally dominant: the 7-point stencil
        # with diagonal +7 / -7 on even / odd vertices (plus a complex part
        # when asked).  LDL^T without pivoting keeps every Schur complement
        # diagonally dominant, so the factor has no element growth and two
        # implementations agree to rounding.
        base = px.Spm.laplacian7(nx, ny, nz, 6.0, 1.0)
        cp, rp, v = base.arrays()
        v = v.astype(np.complex128) if kw.get("complex_") else v.copy()
        for j in range(nx * ny * nz):
            for p in range(cp[j], cp[j + 1]):
                if rp[p] == j:
                    v[p] = (7.0 if j % 2 == 0 else -7.0) + (0.3j if kw.get("complex_") else 0.0)
        return px.Spm.from_csc(nx * ny * nz, cp, rp, v, mtxtype=base.s.mtxtype)
    raise ValueError(kind)


def analyzed(spm, nx, ny=None, nz=None, facto=None, mixed=0, extra=None, dof=1):
    """pastix_data after the grid analysis (example/ordering_grid.c flow)."""
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    iparm, dparm = px.initParam()
    iparm[px.IPARM_ORDERING] = px.PastixOrderPersonal
    iparm[px.IPARM_FACTORIZATION] = px.PastixFactPOTRF if facto is None else facto
    iparm[px.IPARM_MIXED] = mixed
    iparm[px.IPARM_VERBOSE] = px.PastixVerboseNot
    for k, v in (extra or {}).items():
        iparm[k] = v
    pd = px.init(iparm, dparm)
    px.analyze_grid(pd, spm, nx, ny, nz, dof=dof)
    return pd, iparm, dparm


def permute_rhs(pd, b):
    perm = px.order_export(pd)[0]
    bp = np.zeros_like(b)
    bp[perm] = b
    return bp


def unpermute(pd, xp):
    perm = px.order_export(pd)[0]
    return xp[perm]


def static_pivot_threshold(dparm, dtype):
    eps = np.finfo(dtype).eps / 2.0
    return eps * dparm[px.DPARM_A_NORM]


def cblk_diff(solver, La, Lb, dtype, full_diag=False):
    """Largest per-cblk ||La - Lb||_max / (||Lb||_F * eps) over the stored lower
    part (cpucblk_zdiff criterion, kernels/cpucblk_zdiff.c:63-130); with
    full_diag the whole diagonal tile is compared (LU factors)."""
    cb = solver["cblk"].reshape(-1, 6)
    bk = solver["blok"].reshape(-1, 5)
    eps = np.finfo(dtype).eps / 2.0
    worst = 0.0
    for k in range(solver["cblknbr"]):
        n = cb[k, 1] - cb[k, 0] + 1
        off0, off1 = cb[k, 4], cb[k + 1, 4]
        wide = np.complex128 if np.iscomplexobj(Lb) else np.float64
        a = La[off0:off1].astype(wide)
        b = Lb[off0:off1].astype(wide)
        # the diagonal tile: compare its lower triangle only
        d_a = a[:n * n].reshape(n, n).T
        d_b = b[:n * n].reshape(n, n).T
        if not full_diag:
            d_a, d_b = np.tril(d_a), np.tril(d_b)
        diff = max(np.abs(d_a - d_b).max(), np.abs(a[n * n:] - b[n * n:]).max() if off1 - off0 > n * n else 0.0)
        nrm = np.sqrt((np.abs(d_b) ** 2).sum() + (np.abs(b[n * n:]) ** 2).sum())
        if nrm > 0:
            worst = max(worst, diff / (nrm * eps))
    return worst
