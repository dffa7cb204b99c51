"""The learned cost model in the loop: the kernel class of every blok update
(64 x 64 tiles, tall tiles, one wavefront) chosen by the smallest predicted
time (pastix_amd_choose_update_kernels), with the perceptron or with the
per-class polynomial models (the models.c fallback), both fitted to timing
logs of other problems recorded under each class.

- the factor of every choice matches the oracle's factor (a choice changes
  the kernels, never the result beyond rounding);
- the records carry the class that ran them, and a choice can be dropped
  again (the cut-offs return);
- what the network gains: on held-out update records its time predictions
  are closer to the measurements than the per-class polynomials'."""
import numpy as np
import pytest

import pastix_amd as px
import pastix_amd.models as pm
import oracle_ref as orc
from helpers import cblk_diff, static_pivot_threshold

pytestmark = pytest.mark.gpu

TRAIN = [("14^3 Laplacian", "lap7", 14, 1, "POTRF"), ("22^3 Laplacian", "lap7", 22, 1, "POTRF")]
TEST = ("9^3 x 3 dof vector Laplacian", "vlap7", 9, 3, "POTRF")


def test_records_carry_the_update_class():
    rec = pm.class_logs(px, TRAIN[:1], mixed=1, log=lambda m: None)
    upd = rec[rec["ktype"] >= pm.KTYPE["GEMMCblk1d1d"]]
    other = rec[rec["ktype"] < pm.KTYPE["GEMMCblk1d1d"]]
    assert set(np.unique(upd["variant"])) == {0, 1, 2}
    assert (other["variant"] == -1).all()
    # the one-wavefront class only where it fits
    small = upd[upd["variant"] == 2]
    assert (small["m"] <= 256).all() and (small["n"] <= 64).all() and (small["k"] <= 32).all()
    assert (small["m"] > 64).any() and (small["n"] > 16).any()


def test_learned_choice_factor_matches_oracle_and_beats_polynomials():
    out, pd, spm, dparm = pm.choice_experiment(px, TRAIN, TEST, mixed=1, steps=1500, reps=2,
                                               log=lambda m: None)
    ho = out["held_out"]
    assert ho["learned"]["median_rel_err"] <= ho["polynomial"]["median_rel_err"]
    s = px.solver_export(pd)
    F = orc.OracleFactor(s, px.bcsc_export(pd), static_pivot_threshold(dparm, np.float32), np.float32)
    # the last choice of the experiment (the learned one) is in force
    L = px.get_coeftab(pd, s["coefnbr"], np.float32)
    assert cblk_diff(s, L, F.L, np.float32) <= 10.0
    # the polynomial choice, and the cut-offs again
    poly = pm.class_models(pm.class_logs(px, TRAIN[:1], mixed=1, log=lambda m: None), 0)
    for kw in (dict(poly=poly), {}):
        pm.choose_update_kernels(pd, 0, **kw)
        px.task_numfact(pd, spm)
        L = px.get_coeftab(pd, s["coefnbr"], np.float32)
        assert cblk_diff(s, L, F.L, np.float32) <= 10.0
    for name in ("cut-offs", "polynomial", "learned"):
        assert out["choices"][name]["fact_seconds"] > 0
    px.finalize(pd)
