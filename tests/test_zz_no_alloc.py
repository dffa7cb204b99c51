"""The allocation-free solve (pxd_solve_noalloc) against pxd_solve, both
called through the C ABI on the same device vectors.

Two calls in a row, the second with fewer right-hand sides, on double,
single (mixed) and complex factors; then a second, smaller instance sharing
the same scratch; then a problem whose solve needs more scratch than the
static array holds, which is refused with PASTIX_ERR_BADPARAMETER while one
column fewer fits.  The solve kernels add with atomics, so two solves agree
to rounding (the bound is in the factor's precision), not bit for bit."""
import ctypes

import numpy as np
import pytest

import pastix_amd as px
from helpers import analyzed

pytestmark = pytest.mark.gpu

_p = ctypes.c_void_p
for _name in ("pxd_solve", "pxd_solve_noalloc"):
    getattr(px.lib, _name).restype = ctypes.c_int
    getattr(px.lib, _name).argtypes = [_p, ctypes.c_int64, _p, ctypes.c_int64]

NOALLOC_BYTES = 32 << 20   # PXD_NOALLOC_BYTES


def _device(pd):
    dev = _p()
    px._check("device", px.lib.pastix_amd_device_handle(pd, ctypes.byref(dev)))
    return dev


def _solve(fn, dev, b):
    """rc and the result of fn on a device copy of b (n x nrhs, ld n)."""
    b = np.asfortranarray(b)
    n, nrhs = b.shape
    d = _p()
    px._check("pxd_malloc", px.lib.pxd_malloc(dev, b.nbytes, ctypes.byref(d)))
    try:
        px.lib.pxd_h2d(dev, d, b.ctypes.data_as(_p), b.nbytes)
        rc = fn(dev, nrhs, d, n)
        out = np.zeros_like(b)
        px.lib.pxd_d2h(dev, out.ctypes.data_as(_p), d, b.nbytes)
    finally:
        px.lib.pxd_free(dev, d)
    return rc, out


def _compare(dev, spm, b, eps):
    r1, xa = _solve(px.lib.pxd_solve, dev, b)
    r2, xb = _solve(px.lib.pxd_solve_noalloc, dev, b)
    assert r1 == 0 and r2 == 0
    scale = np.abs(xa).max()
    assert np.abs(xb - xa).max() <= 64 * eps * scale


@pytest.mark.parametrize("kind,facto,mixed", [("lap7", None, 0), ("lap7", None, 1),
                                              ("helm", px.PastixFactLDLT, 0), ("conv", px.PastixFactGETRF, 1)])
def test_noalloc_solve_matches_pxd_solve(kind, facto, mixed):
    nx = 11
    if kind == "lap7":
        spm = px.Spm.laplacian7(nx, nx, nx)
    elif kind == "helm":
        spm = px.Spm.helmholtz7(nx, nx, nx, 0.5, 0.1)
    else:
        spm = px.Spm.convdiff27(nx, nx, nx, 1.0, 0.5, 0)
    pd, iparm, dparm = analyzed(spm, nx, facto=facto, mixed=mixed)
    px.task_numfact(pd, spm)
    dev = _device(pd)
    eps = np.finfo(np.float32 if mixed else np.float64).eps
    for nrhs, seed in ((11, 0), (3, 1)):   # more than one pass, then a smaller input
        _, b = spm.gen_rhs(nrhs, seed=seed)
        _compare(dev, spm, b, eps)
    px.finalize(pd)
    # another, smaller instance on the same GPU (the scratch is shared)
    spm2 = px.Spm.laplacian7(7, 6, 5)
    pd2, _, _ = analyzed(spm2, 7, 6, 5)
    px.task_numfact(pd2, spm2)
    _, b2 = spm2.gen_rhs(2, seed=2)
    _compare(_device(pd2), spm2, b2, np.finfo(np.float64).eps)
    px.finalize(pd2)


def test_noalloc_solve_refuses_beyond_its_scratch():
    nx, ny = 800, 700      # 560,000 unknowns: 8 double columns need 35.8 MB > 32 MiB
    spm = px.Spm.laplacian7(nx, ny, 1)
    pd, _, _ = analyzed(spm, nx, ny, 1)
    px.task_numfact(pd, spm)
    dev = _device(pd)
    n = spm.n
    assert n * 8 * 8 > NOALLOC_BYTES >= n * 7 * 8
    _, b = spm.gen_rhs(8, seed=0)
    rc, _ = _solve(px.lib.pxd_solve_noalloc, dev, b)
    assert rc == 7
    _compare(dev, spm, np.asfortranarray(b[:, :7]), np.finfo(np.float64).eps)
    px.finalize(pd)
