"""Kernel timing log and learned cost model on the GPU.

The timing log must hold one record per panel, per TRSM of a cblk and per
blok update with the sizes the reference's kernel_trace_stop would record
(kernels/core_zpotrfsp.c:258, kernels/core_ztrsmsp.c:366-397,
kernels/gpu_zgemmsp.c:242-251), computed here from the exported solver
matrix; the factorization that writes it (one launch per blok update) must
still match the oracle.  The perceptron's loss, gradient and forward pass are
checked against the float64 restatement in oracle_ref.
"""
import numpy as np
import pytest

import pastix_amd as px
import pastix_amd.models as pm
import oracle_ref as orc
from helpers import analyzed, cblk_diff, make_problem, static_pivot_threshold

pytestmark = pytest.mark.gpu
K = pm.KTYPE


def expected_records(s, lu):
    """Sorted (ktype, m, n, k) of every logged task, from the solver matrix."""
    cb = s["cblk"].reshape(-1, 6)
    bk = s["blok"].reshape(-1, 5)
    out = []
    panel = K["GETRF"] if lu else K["POTRF"]
    for c in range(s["cblknbr"]):
        n = int(cb[c, 1] - cb[c, 0] + 1)
        stride = int(cb[c, 3])
        out.append((panel, n, 0, 0))
        if stride > n:
            out += [(K["TRSMCblk2d"], stride - n, n, 0)] * (2 if lu else 1)
        for b in range(cb[c, 2] + 1, cb[c + 1, 2]):
            rowoff = int(bk[b, 4] // n)
            N = int(bk[b, 1] - bk[b, 0] + 1)
            out.append((K["GEMMCblk2d2d"], stride - rowoff, N, n))
            if lu and stride - rowoff - N > 0:
                out.append((K["GEMMCblk2d2d"], stride - rowoff - N, N, n))
    return sorted(out)


@pytest.mark.parametrize("facto,dtype", [(None, np.float64), (None, np.float32),
                                         (px.PastixFactGETRF, np.float64)])
def test_timing_log_records_and_factor(facto, dtype, tmp_path):
    nx = 12
    kind = "convdiff27" if facto == px.PastixFactGETRF else "lap7"
    spm = make_problem(kind, nx)
    mixed = int(dtype == np.float32)
    pd, iparm, dparm = analyzed(spm, nx, facto=facto, mixed=mixed)
    px.set_option(pd, "model_log", 1)
    px.task_numfact(pd, spm)
    s = px.solver_export(pd)
    lu = facto == px.PastixFactGETRF
    rec = pm.timing_log(pd)
    got = sorted(zip(rec["ktype"].tolist(), rec["m"].tolist(), rec["n"].tolist(), rec["k"].tolist()))
    assert got == expected_records(s, lu)
    assert (rec["time"] > 0).all() and (rec["time"] < 0.1).all()
    # the per-update launches give the same factor
    L = px.get_coeftab(pd, s["coefnbr"], dtype)
    if lu:
        F = orc.OracleLU(s, px.bcsc_export(pd), static_pivot_threshold(dparm, dtype), dtype)
        U = px.get_coeftab(pd, s["coefnbr"], dtype, side=1)
        assert cblk_diff(s, L, F.L, dtype, full_diag=True) <= 10.0
        assert cblk_diff(s, U, F.U, dtype, full_diag=True) <= 10.0
    else:
        F = orc.OracleFactor(s, px.bcsc_export(pd), static_pivot_threshold(dparm, dtype), dtype)
        assert cblk_diff(s, L, F.L, dtype) <= 10.0
    # model.csv round trip and a polynomial fit of the log
    f = str(tmp_path / "model.csv")
    pm.write_model_csv(pd, f)
    back = pm.read_model_csv(f)
    for fld in ("ktype", "m", "n", "k"):
        assert np.array_equal(back[fld], rec[fld])
    assert np.allclose(back["time"], rec["time"], rtol=1e-5)
    m = pm.Model().fit(rec, 0 if mixed else 1)
    assert (m.costs(0 if mixed else 1, rec) >= 0).all()
    # a factorization without the log leaves the previous records alone and
    # still matches the oracle
    px.set_option(pd, "model_log", 0)
    px.task_numfact(pd, spm)
    assert len(pm.timing_log(pd)) == len(rec)
    px.finalize(pd)


def _mlp_case(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(0, 1, (n, orc.MLP_F)).astype(np.float32)
    t = (np.sin(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2] - 0.3 * X[:, 3]).astype(np.float32)
    return X, t


@pytest.mark.parametrize("n", [1, 63, 64, 1000])
def test_mlp_loss_and_gradient_match_oracle(n):
    X, t = _mlp_case(n, 3)
    net = pm.Mlp(seed=11)
    p = net.params()
    net.set_data(X, t)
    loss, g = net.loss_grad()
    lo, go = orc.mlp_loss_grad(p, X.astype(np.float64), t.astype(np.float64))
    assert abs(loss - lo) <= 1e-5 * max(1.0, lo)
    assert np.abs(g - go).max() <= 1e-5 * max(1.0, np.abs(go).max())
    # deterministic reduction: the same data gives the same bits
    loss2, g2 = net.loss_grad()
    assert loss2 == loss and np.array_equal(g, g2)
    y = net.predict(X)
    assert np.allclose(y, orc.mlp_forward(p, X.astype(np.float64)), rtol=1e-5, atol=1e-5)


def test_mlp_training_and_set_params():
    X, t = _mlp_case(4096, 5)
    net = pm.Mlp(seed=2)
    p0 = net.params()
    net.set_data(X, t)
    l0, _ = net.loss_grad()
    l1 = net.train(300, 1e-2)
    assert l1 < 0.2 * l0
    net.set_params(p0)
    assert np.array_equal(net.params(), p0)
    assert net.loss_grad()[0] == l0


def test_learned_model_on_a_synthetic_log():
    rng = np.random.default_rng(0)
    n = 3000
    rec = np.zeros(n, dtype=pm.ENTRY)
    rec["ktype"] = K["GEMMCblk2d2d"]
    rec["m"] = rng.integers(1, 2000, n)
    rec["n"] = rng.integers(1, 320, n)
    rec["k"] = rng.integers(1, 320, n)
    flops = 2.0 * rec["m"] * rec["n"] * rec["k"]
    rec["time"] = 3e-6 + flops / 50e12 * rng.lognormal(0, 0.05, n)
    out, poly, learned = pm.compare(rec, 1, steps=1500)
    assert out["update"]["learned"]["median_rel_err"] < 0.1
    assert out["update"]["polynomial"]["median_rel_err"] < 0.2
