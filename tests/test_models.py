"""Kernel cost models (CPU tests): the polynomial forms, the kernel names, the
propagation rules and defaults, the model-file reader and writer, the
least-squares fit, and the float64 restatement of the learned model's
gradient.

Expected values are the reference's documented formulas evaluated at sizes
chosen so that every monomial has a distinct value (a swapped coefficient
index fails), the default tables of common/models.c:337-474 and hand
derivations of the propagation rules of common/models.c:100-201.
"""
import os

import numpy as np
import pytest

import pastix_amd as px
import pastix_amd.models as pm
import oracle_ref as orc

K = pm.KTYPE
M_, N_, K_ = 3, 5, 7


def unit(i, n=8):
    c = np.zeros(n)
    c[i] = 1.0
    return c


def test_cost3_terms_follow_the_reference_formula():
    # a7 MNK + a6 MK + a5 KN + a4 MN + a3 M + a2 N + a1 K + a0 (common/models.h:74-81)
    expect = [1, K_, N_, M_, M_ * N_, K_ * N_, M_ * K_, M_ * N_ * K_]
    assert [pm.cost3(unit(i), M_, N_, K_) for i in range(8)] == expect
    assert pm.cost3(-unit(7), M_, N_, K_) == 0.0          # clamped at zero


def test_cost2_and_cost1_terms():
    # a5 MN^2 + a4 MN + a3 N^2 + a2 M + a1 N + a0 (common/models.h:57-58)
    assert [pm.cost2(unit(i, 6), M_, N_) for i in range(6)] == [1, N_, M_, N_ * N_, M_ * N_, M_ * N_ * N_]
    # a3 N^3 + a2 N^2 + a1 N + a0 (common/models.h:42-43)
    assert [pm.cost1(unit(i, 4), N_) for i in range(4)] == [1, N_, N_ ** 2, N_ ** 3]
    assert pm.cost1([-1.0, 0, 0, 0], 4) == 0.0


def test_kernel_names():
    assert pm.kernel_id("potrf") == (K["POTRF"], 4)
    assert pm.kernel_id("GETRF") == (K["GETRF"], 4)            # strcasecmp
    assert pm.kernel_id("trsmblok2d") == (K["TRSMBlok2d"], 6)
    assert pm.kernel_id("gemmcblk2d2d") == (K["GEMMCblk2d2d"], 8)
    assert pm.kernel_id("gemmbloklrlr") == (K["GEMMBlokLRLR"], 8)
    assert pm.kernel_id("scalocblk") == (-1, 0)                # no model-file name
    assert pm.kernel_id("nope") == (-1, 0)
    assert pm.NKTYPES == 22 and K["GEMMCblk2d2d"] == 14 and K["TRSMCblk2d"] == 8


def test_default_cpu_model_and_propagation():
    m = pm.Model.default_cpu()
    assert m.name == "AMD Opteron 6180 - Intel MKL"
    c = m.coefficients
    potrf = np.array([4.071507e-07, -1.469893e-07, 1.707006e-08, 2.439599e-11])
    assert np.array_equal(c[1, K["POTRF"], :4], potrf)
    # GETRF: twice the quadratic and cubic terms; the other panels: copies
    assert np.array_equal(c[1, K["GETRF"], :4], potrf * [1, 1, 2, 2])
    for k in ("HETRF", "PXTRF", "SYTRF"):
        assert np.array_equal(c[1, K[k], :4], potrf)
    # other arithmetics: (a+1)/2 of double (s 0.5, c 1.5, z 2)
    for a, r in ((0, 0.5), (2, 1.5), (3, 2.0)):
        assert np.allclose(c[a, K["POTRF"], :4], r * potrf, rtol=1e-15)
        assert np.allclose(c[a, K["GETRF"], :4], r * potrf * [1, 1, 2, 2], rtol=1e-15)
    trsm = np.array([3.255168e-06, 3.976198e-08, 0, 0, 0, 2.626177e-10, 0, 0])
    for k in ("TRSMCblk1d", "TRSMCblk2d", "TRSMCblkLR", "TRSMBlok2d"):
        assert np.array_equal(c[1, K[k]], trsm)
        assert np.allclose(c[3, K[k]], 2 * trsm, rtol=1e-15)
    assert c[1, K["TRSMBlokLR"], 0] == pm.UNSET           # outside the TRSM family range
    gemm = np.array([1.216278e-06, 0., -2.704179e-10, 1.148989e-07, 2.724804e-10, 1.328900e-09, 0.,
                     2.429169e-10])
    for k in ("GEMMCblk1d1d", "GEMMCblk1d2d", "GEMMCblk2d2d", "GEMMCblkFRLR", "GEMMCblkLRLR"):
        assert np.array_equal(c[1, K[k]], gemm)
    assert np.array_equal(c[1, K["GEMMBlok2d2d"]], [0] * 7 + [2. / 24.e9])
    # the block GEMM line comes after the panel GEMM had already filled the
    # other arithmetics, so there they keep the panel GEMM's scaled values
    assert np.allclose(c[0, K["GEMMBlok2d2d"]], 0.5 * gemm, rtol=1e-15)
    assert np.allclose(c[2, K["GEMMBlok2d2d"]], 1.5 * gemm, rtol=1e-15)
    for k in ("SCALOCblk", "SCALOBlok", "GEMMBlokLRLR", "GEADDCblkFRFR"):
        assert c[1, K[k], 0] == pm.UNSET
    assert m.cost(1, K["SCALOCblk"], 10) == -1.0
    # the cost dispatch uses the kernel family's form
    assert m.cost(1, K["POTRF"], 64) == pm.cost1(potrf, 64)
    assert m.cost(1, K["TRSMCblk2d"], 100, 64) == pm.cost2(trsm, 100, 64)
    assert m.cost(1, K["GEMMCblk2d2d"], 100, 64, 32) == pm.cost3(gemm, 100, 64, 32)


def test_default_gpu_model():
    m = pm.Model.default_gpu()
    assert m.name == "Nvidia K40 GK1108L - CUDA 8.0"
    c = m.coefficients
    tb = [-3.16663635648446e-05, 2.63809317549331e-06, 5.86447245256688e-07, -1.57859559108480e-09,
          -4.74303242824929e-09, 5.36284121953867e-12, 0, 0]
    for k in ("TRSMCblk1d", "TRSMCblk2d", "TRSMCblkLR", "TRSMBlok2d"):
        assert np.array_equal(c[1, K[k]], tb)
    assert c[1, K["POTRF"], 0] == pm.UNSET and c[0, K["GETRF"], 0] == pm.UNSET
    assert np.array_equal(c[1, K["GEMMBlok2d2d"]], [0] * 7 + [2. / 1.2e12])


def test_model_file_reader(tmp_path):
    f = tmp_path / "model.txt"
    f.write_text("# comment\n# another\nMy unit\n"
                 "0;potrf;1e-6;2e-7;3e-9;4e-12\n"
                 "# skipped\n"
                 "1;gemmcblk2d2d;1;2;3;4;5;6;7;8\n"
                 "1;trsmcblk2d;1;2;3\n"          # too few coefficients: skipped
                 "7;potrf;1;2;3;4\n"              # bad arithmetic: skipped
                 "1;nokernel;1;2;3;4\n")          # unknown kernel: skipped
    m = pm.Model.read(str(f))
    assert m.name == "My unit"
    c = m.coefficients
    p0 = np.array([1e-6, 2e-7, 3e-9, 4e-12])
    assert np.array_equal(c[0, K["POTRF"], :4], p0)
    assert np.array_equal(c[0, K["GETRF"], :4], p0 * [1, 1, 2, 2])
    for a in (1, 2, 3):       # ratio (a+1)/(0+1)
        assert np.allclose(c[a, K["POTRF"], :4], (a + 1) * p0, rtol=1e-15)
    g = np.arange(1.0, 9.0)
    assert np.array_equal(c[1, K["GEMMCblk2d2d"]], g)
    assert np.array_equal(c[1, K["GEMMBlok2d2d"]], g)
    assert np.allclose(c[3, K["GEMMCblk1d1d"]], 2 * g)
    assert c[1, K["TRSMCblk2d"], 0] == pm.UNSET
    with pytest.raises(OSError):
        pm.Model.read(str(tmp_path / "missing.txt"))


def _synthetic_records(rng, n=400):
    ms = rng.integers(1, 600, n)
    ns = rng.integers(1, 320, n)
    ks = rng.integers(1, 320, n)
    rec = np.zeros(3 * n, dtype=pm.ENTRY)
    gemm = np.array([2e-6, 1e-9, 2e-9, 3e-9, 4e-11, 5e-11, 6e-11, 2e-13])
    trsm = np.array([3e-6, 1e-9, 2e-9, 3e-11, 4e-11, 5e-13])
    potrf = np.array([1e-6, 1e-8, 1e-10, 1e-12])
    for i in range(n):
        rec[i] = (K["GEMMCblk2d2d"], ms[i], ns[i], ks[i], pm.cost3(gemm, ms[i], ns[i], ks[i]), 0, 0)
        rec[n + i] = (K["TRSMCblk2d"], ms[i], ns[i], 0, pm.cost2(trsm, ms[i], ns[i]), -1, 0)
        rec[2 * n + i] = (K["POTRF"], ns[i], 0, 0, pm.cost1(potrf, ns[i]), -1, 0)
    return rec, gemm, trsm, potrf


def test_fit_recovers_exact_polynomials(tmp_path):
    rng = np.random.default_rng(0)
    rec, gemm, trsm, potrf = _synthetic_records(rng)
    m = pm.Model().fit(rec, 1, name="synthetic")
    c = m.coefficients
    assert m.name == "synthetic"
    assert np.allclose(c[1, K["GEMMCblk2d2d"]], gemm, rtol=1e-6, atol=0)
    assert np.allclose(c[1, K["TRSMCblk2d"], :6], trsm, rtol=1e-6, atol=0)
    assert np.allclose(c[1, K["POTRF"], :4], potrf, rtol=1e-6, atol=0)
    # propagated like a model file line
    assert np.allclose(c[0, K["GEMMCblk2d2d"]], 0.5 * c[1, K["GEMMCblk2d2d"]], rtol=1e-15)
    # the written file reads back bit for bit
    f = str(tmp_path / "fit.txt")
    m.write(f)
    r = pm.Model.read(f)
    assert r.name == "synthetic"
    assert np.array_equal(r.coefficients, c)


def test_instance_models_from_environment(tmp_path, monkeypatch):
    f = tmp_path / "cpu.txt"
    f.write_text("Host model\n1;potrf;1;2;3;4\n")
    monkeypatch.setenv("PASTIX_MODELS_CPU", str(f))
    monkeypatch.delenv("PASTIX_MODELS_GPU", raising=False)
    iparm, dparm = px.initParam()
    pd = px.init(iparm, dparm)
    assert pm.Model.of_instance(pd, gpu=False).name == "Host model"
    assert pm.Model.of_instance(pd, gpu=True).name == "Nvidia K40 GK1108L - CUDA 8.0"
    # no factorization ran: the timing log is empty and model.csv has its header only
    assert len(pm.timing_log(pd)) == 0
    out = tmp_path / "model.csv"
    pm.write_model_csv(pd, str(out))
    assert out.read_text().splitlines()[2] == "Kernel;M;N;K;Time"
    assert len(pm.read_model_csv(str(out))) == 0
    px.finalize(pd)


def test_mlp_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    npar = orc.MLP_F * orc.MLP_H + orc.MLP_H + orc.MLP_H ** 2 + 2 * orc.MLP_H + 1
    p = rng.normal(0, 0.3, npar)
    X = rng.normal(0, 1, (40, orc.MLP_F))
    t = rng.normal(0, 1, 40)
    loss, g = orc.mlp_loss_grad(p, X, t)
    assert np.isclose(loss, np.mean((orc.mlp_forward(p, X) - t) ** 2), rtol=1e-14)
    h = 1e-6
    for i in rng.choice(npar, 60, replace=False).tolist() + [npar - 1, 0]:
        e = np.zeros(npar)
        e[i] = h
        fd = (orc.mlp_loss_grad(p + e, X, t)[0] - orc.mlp_loss_grad(p - e, X, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 * max(1.0, abs(g[i])), i
