"""Kernel cost models: the reference's polynomial models and model files,
the engine's kernel timing log, and the learned (GPU MLP) model.

Mirrors ``common/models.[ch]`` of the reference (``modelsGetCost{1,2,3}Param``,
``modelsRead``, ``modelsPropagate``, ``modelsInitDefault{CPU,GPU}``,
``pastixModels{New,Free,Load}``) through the C-ABI library, and adds:

- ``timing_log(pastix_data)``: the records of the last factorization run with
  the option ``model_log`` = 1 (numpy structured array ktype, m, n, k, time),
  ``write_model_csv`` writing them in the reference's model.csv format
  (``kernels/kernels_trace.c:222-249``);
- ``Model.fit(records, arith)``: least-squares fit (relative error) of the
  polynomials, ``Model.write`` writing a model file ``modelsRead`` reads;
- ``LearnedModel``: the GPU multilayer perceptron fitted to log times.

``python -m pastix_amd.models --nx 100`` factorizes the 7-point Laplacian with
the timing log on, fits both models on a random 80 % of the records and
prints the error of each on the other 20 % (needs the GPU).

``python -m pastix_amd.models --transfer`` logs the factorizations of the
20^3 and 100^3 Laplacians and of the 64^3 convection-diffusion LU, fits both
models on those records, and scores them, next to the reference's built-in
defaults (the fallback of ``common/models.c`` when no model file is given),
on the log of the 60^3 x 3-dof vector Laplacian, a problem none of the
training runs contains.
"""
import argparse
import ctypes
import json

import numpy as np

from . import lib, _check, _ptr, _sig, pastix_int, _p, _I

KTYPES = ["GETRF", "HETRF", "POTRF", "PXTRF", "SYTRF", "SCALOCblk", "SCALOBlok", "TRSMCblk1d",
          "TRSMCblk2d", "TRSMCblkLR", "TRSMBlok2d", "TRSMBlokLR", "GEMMCblk1d1d", "GEMMCblk1d2d",
          "GEMMCblk2d2d", "GEMMCblkFRLR", "GEMMCblkLRLR", "GEMMBlok2d2d", "GEMMBlokLRLR",
          "GEADDCblkFRFR", "GEADDCblkFRLR", "GEADDCblkLRLR"]
KTYPE = {name: i for i, name in enumerate(KTYPES)}
NKTYPES = len(KTYPES)
UNSET = 3735928559.0

# pxd_model_entry_t; variant: kernel class of a blok update (0: 64 x 64
# tiles, 1: tall tiles, 2: one wavefront), -1 for other records
ENTRY = np.dtype([("ktype", np.int32), ("m", np.int32), ("n", np.int32), ("k", np.int32),
                  ("time", np.float64), ("variant", np.int32), ("pad", np.int32)])
MLP_INPUTS = 10   # PXD_MLP_F


class pastix_model_t(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p),
                ("coefficients", ctypes.c_double * 8 * NKTYPES * 4)]


_PM = ctypes.POINTER(pastix_model_t)
_D = ctypes.c_double
_sig("modelsGetCost1Param", _D, _p, pastix_int)
_sig("modelsGetCost2Param", _D, _p, pastix_int, pastix_int)
_sig("modelsGetCost3Param", _D, _p, pastix_int, pastix_int, pastix_int)
_sig("modelsGetKernelId", _I, ctypes.c_char_p, ctypes.POINTER(_I))
_sig("modelsPropagate", None, _PM, _I, _I)
_sig("modelsRead", _I, _PM, ctypes.c_char_p)
_sig("modelsInitDefaultCPU", _I, _PM)
_sig("modelsInitDefaultGPU", _I, _PM)
_sig("pastixModelsNew", _PM)
_sig("pastixModelsFree", None, _PM)
_sig("pastix_amd_models_get", _PM, _p, _I)
_sig("pastix_amd_model_cost", _D, _PM, _I, _I, pastix_int, pastix_int, pastix_int)
_sig("pastix_amd_model_log", _I, _p, _p, pastix_int, ctypes.POINTER(pastix_int))
_sig("pastix_amd_model_write_csv", _I, _p, ctypes.c_char_p)
_sig("pastix_amd_models_fit", _I, _PM, ctypes.c_char_p, _I, _p, pastix_int)
_sig("pastix_amd_models_write", _I, _PM, ctypes.c_char_p)
_sig("pastix_amd_learned_train", _I, ctypes.POINTER(_p), _I, _I, _p, pastix_int, _I, _D,
     ctypes.c_uint64, ctypes.POINTER(_D))
_sig("pastix_amd_learned_predict", _I, _p, _I, _p, pastix_int, _p)
_sig("pastix_amd_learned_free", None, _p)
_sig("pxd_mlp_nparams", _I)
_sig("pxd_mlp_create", _I, ctypes.POINTER(_p), _I, ctypes.c_uint64)
_sig("pxd_mlp_destroy", None, _p)
_sig("pxd_mlp_get_params", _I, _p, _p)
_sig("pxd_mlp_set_params", _I, _p, _p)
_sig("pxd_mlp_set_data", _I, _p, _p, _p, ctypes.c_int64)
_sig("pxd_mlp_loss_grad", _I, _p, ctypes.POINTER(_D), _p)
_sig("pxd_mlp_train", _I, _p, _I, _D, ctypes.POINTER(_D))
_sig("pxd_mlp_predict", _I, _p, _p, ctypes.c_int64, _p)
_sig("pastix_amd_choose_update_kernels", _I, _p, _p, _p, _I, ctypes.POINTER(_D))


def cost1(coefs, n):
    c = np.ascontiguousarray(coefs, dtype=np.float64)
    return lib.modelsGetCost1Param(_ptr(c), n)


def cost2(coefs, m, n):
    c = np.ascontiguousarray(coefs, dtype=np.float64)
    return lib.modelsGetCost2Param(_ptr(c), m, n)


def cost3(coefs, m, n, k):
    c = np.ascontiguousarray(coefs, dtype=np.float64)
    return lib.modelsGetCost3Param(_ptr(c), m, n, k)


def kernel_id(name):
    nb = _I()
    kid = lib.modelsGetKernelId(name.encode(), ctypes.byref(nb))
    return kid, nb.value


class Model:
    """A pastix_model_t owned by Python (or borrowed from an instance)."""

    def __init__(self, ptr=None, owned=True):
        self.ptr = ptr if ptr is not None else lib.pastixModelsNew()
        self.owned = owned and ptr is None

    def __del__(self):
        if getattr(self, "owned", False) and self.ptr:
            lib.pastixModelsFree(self.ptr)
            self.ptr = None

    @classmethod
    def default_cpu(cls):
        m = cls()
        lib.modelsInitDefaultCPU(m.ptr)
        return m

    @classmethod
    def default_gpu(cls):
        m = cls()
        lib.modelsInitDefaultGPU(m.ptr)
        return m

    @classmethod
    def read(cls, filename):
        m = cls()
        if lib.modelsRead(m.ptr, filename.encode()) != 0:
            raise OSError("modelsRead failed for %s" % filename)
        return m

    @classmethod
    def of_instance(cls, pastix_data, gpu):
        p = lib.pastix_amd_models_get(pastix_data, 1 if gpu else 0)
        return cls(p, owned=False) if p else None

    @property
    def name(self):
        n = self.ptr.contents.name
        return n.decode() if n else None

    @property
    def coefficients(self):
        """Copy as a (4, NKTYPES, 8) array (UNSET marks unset kernels)."""
        return np.ctypeslib.as_array(self.ptr.contents.coefficients).copy()

    def set(self, arith, ktype, coefs):
        row = self.ptr.contents.coefficients[arith][ktype]
        for i, v in enumerate(coefs):
            row[i] = v

    def propagate(self, arith, ktype):
        lib.modelsPropagate(self.ptr, arith, ktype)

    def cost(self, arith, ktype, m, n=0, k=0):
        return lib.pastix_amd_model_cost(self.ptr, arith, ktype, m, n, k)

    def costs(self, arith, records):
        return np.array([self.cost(arith, int(r["ktype"]), int(r["m"]), int(r["n"]), int(r["k"]))
                         for r in records])

    def fit(self, records, arith, name=None):
        rec = np.ascontiguousarray(records, dtype=ENTRY)
        _check("pastix_amd_models_fit", lib.pastix_amd_models_fit(
            self.ptr, name.encode() if name else None, arith, _ptr(rec), len(rec)))
        return self

    def write(self, filename):
        _check("pastix_amd_models_write", lib.pastix_amd_models_write(self.ptr, filename.encode()))


def timing_log(pastix_data):
    """Records of the last factorization's kernel timing log."""
    cnt = pastix_int()
    _check("pastix_amd_model_log", lib.pastix_amd_model_log(pastix_data, None, 0, ctypes.byref(cnt)))
    out = np.zeros(cnt.value, dtype=ENTRY)
    if cnt.value:
        _check("pastix_amd_model_log", lib.pastix_amd_model_log(pastix_data, _ptr(out), cnt.value,
                                                                ctypes.byref(cnt)))
    return out


def write_model_csv(pastix_data, filename):
    _check("pastix_amd_model_write_csv", lib.pastix_amd_model_write_csv(pastix_data, filename.encode()))


def read_model_csv(filename):
    """Records of a model.csv file (comment lines, then Kernel;M;N;K;Time)."""
    rows = []
    with open(filename) as f:
        for line in f:
            if line.startswith("#") or line.startswith("Kernel"):
                continue
            p = line.strip().split(";")
            if len(p) == 5:
                rows.append((int(p[0]), int(p[1]), int(p[2]), int(p[3]), float(p[4]), -1, 0))
    return np.array(rows, dtype=ENTRY)


class Mlp:
    """The raw GPU perceptron (pxd_mlp_*): parameters, loss and gradient."""

    def __init__(self, device=0, seed=1):
        self.h = _p()
        _check("pxd_mlp_create", lib.pxd_mlp_create(ctypes.byref(self.h), device, seed))
        self.nparams = lib.pxd_mlp_nparams()

    def __del__(self):
        if getattr(self, "h", None):
            lib.pxd_mlp_destroy(self.h)
            self.h = None

    def params(self):
        p = np.zeros(self.nparams, dtype=np.float32)
        _check("pxd_mlp_get_params", lib.pxd_mlp_get_params(self.h, _ptr(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparams
        _check("pxd_mlp_set_params", lib.pxd_mlp_set_params(self.h, _ptr(p)))

    def set_data(self, X, t):
        X = np.ascontiguousarray(X, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        assert X.ndim == 2 and X.shape[1] == MLP_INPUTS and t.shape == (X.shape[0],)
        self._X, self._t = X, t
        _check("pxd_mlp_set_data", lib.pxd_mlp_set_data(self.h, _ptr(X), _ptr(t), X.shape[0]))

    def loss_grad(self):
        loss = _D()
        g = np.zeros(self.nparams, dtype=np.float32)
        _check("pxd_mlp_loss_grad", lib.pxd_mlp_loss_grad(self.h, ctypes.byref(loss), _ptr(g)))
        return loss.value, g

    def train(self, steps, lr):
        loss = _D()
        _check("pxd_mlp_train", lib.pxd_mlp_train(self.h, steps, lr, ctypes.byref(loss)))
        return loss.value

    def predict(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.zeros(X.shape[0], dtype=np.float32)
        _check("pxd_mlp_predict", lib.pxd_mlp_predict(self.h, _ptr(X), X.shape[0], _ptr(y)))
        return y


class LearnedModel:
    """pastix_amd_learned_*: the perceptron fitted to timing-log records."""

    def __init__(self, records, arith, steps=2000, lr=1e-2, seed=1, device=0):
        rec = np.ascontiguousarray(records, dtype=ENTRY)
        self.h = _p()
        loss = _D()
        _check("pastix_amd_learned_train", lib.pastix_amd_learned_train(
            ctypes.byref(self.h), device, arith, _ptr(rec), len(rec), steps, lr, seed, ctypes.byref(loss)))
        self.loss = loss.value
        self.arith = arith

    def __del__(self):
        if getattr(self, "h", None):
            lib.pastix_amd_learned_free(self.h)
            self.h = None

    def predict(self, records):
        rec = np.ascontiguousarray(records, dtype=ENTRY)
        t = np.zeros(len(rec))
        _check("pastix_amd_learned_predict", lib.pastix_amd_learned_predict(
            self.h, self.arith, _ptr(rec), len(rec), _ptr(t)))
        return t


VARIANTS = ("tiles64", "tall", "wavefront")   # kernel classes of a blok update


def choose_update_kernels(pastix_data, arith, learned=None, poly=None):
    """Kernel class of every blok update by the smallest predicted time
    (pastix_amd_choose_update_kernels): with a LearnedModel, else with three
    Models (one per class, fitted to that class's update records), else the
    engine's cut-offs again.  Returns the predicted total update time of the
    choice (0 for the cut-offs).  Applies from the next factorization."""
    pred = _D()
    arr = None
    if learned is None and poly is not None:
        arr = (_PM * 3)(*[m.ptr for m in poly])
    _check("pastix_amd_choose_update_kernels", lib.pastix_amd_choose_update_kernels(
        pastix_data, learned.h if learned is not None else None,
        ctypes.cast(arr, _p) if arr is not None else None, arith, ctypes.byref(pred)))
    return pred.value


def class_models(records, arith):
    """Three polynomial Models of the update time, one per kernel class, each
    fitted to the update records that ran on that class."""
    upd = records[records["ktype"] >= KTYPE["GEMMCblk1d1d"]]
    out = []
    for v in range(3):
        sel = upd[upd["variant"] == v]
        if len(sel) < 8:
            raise ValueError("too few update records of class %s to fit its model" % VARIANTS[v])
        out.append(Model().fit(sel, arith, name=VARIANTS[v]))
    return out


def error_summary(pred, meas):
    """Relative errors |pred - meas| / meas: median, 90th percentile, and the
    ratio of the predicted to the measured total."""
    r = np.abs(pred - meas) / meas
    return dict(median_rel_err=float(np.median(r)), p90_rel_err=float(np.percentile(r, 90)),
                total_ratio=float(pred.sum() / meas.sum()), records=int(len(meas)))


def compare(records, arith, steps=2000, lr=1e-2, seed=7, frac=0.8):
    """Fit the polynomials and the perceptron on a random share of the
    records and compare both with the measured times of the others, per
    kernel family and overall."""
    rng = np.random.default_rng(seed)
    rec = records[records["time"] > 0]
    idx = rng.permutation(len(rec))
    ntr = int(frac * len(rec))
    train, test = rec[idx[:ntr]], rec[idx[ntr:]]
    poly = Model().fit(train, arith, name="fit")
    learned = LearnedModel(train, arith, steps=steps, lr=lr, seed=seed)
    out = dict(train_records=int(len(train)), test_records=int(len(test)), mlp_train_loss=learned.loss)
    pp, pl = poly.costs(arith, test), learned.predict(test)
    fams = dict(panel=test["ktype"] <= KTYPE["SYTRF"],
                trsm=(test["ktype"] >= KTYPE["TRSMCblk1d"]) & (test["ktype"] < KTYPE["GEMMCblk1d1d"]),
                update=test["ktype"] >= KTYPE["GEMMCblk1d1d"])
    fams["all"] = np.ones(len(test), dtype=bool)
    for f, sel in fams.items():
        if sel.any():
            out[f] = dict(measured_total_s=float(test["time"][sel].sum()),
                          polynomial=error_summary(pp[sel], test["time"][sel]),
                          learned=error_summary(pl[sel], test["time"][sel]))
    return out, poly, learned


# Problems of the transfer experiment: (label, generator, grid edge, dof,
# factorization, factor precisions logged)
TRAIN_RUNS = [("20^3 Laplacian POTRF", "lap7", 20, 1, "POTRF", (1,)),
              ("100^3 Laplacian POTRF", "lap7", 100, 1, "POTRF", (0, 1)),
              ("64^3 convection-diffusion GETRF", "convdiff27", 64, 1, "GETRF", (0, 1))]
TEST_RUN = ("60^3 x 3 dof vector Laplacian POTRF", "vlap7", 60, 3, "POTRF", (0, 1))


def _logged_factorization(px, kind, nx, dof, facto, mixed, opts=None):
    """Timing-log records of one factorization (model_log = 1), with the
    engine options opts."""
    if kind == "lap7":
        spm = px.Spm.laplacian7(nx, nx, nx, 6.0, 1.0)
    elif kind == "convdiff27":
        spm = px.Spm.convdiff27(nx, nx, nx, 1.0, 0.5, 0)
    else:
        spm = px.Spm.vlaplacian7(nx, nx, nx, dof, 1.0, 0)
    iparm, dparm = px.initParam()
    iparm[px.IPARM_ORDERING] = px.PastixOrderPersonal
    iparm[px.IPARM_FACTORIZATION] = px.PastixFactGETRF if facto == "GETRF" else px.PastixFactPOTRF
    iparm[px.IPARM_MIXED] = mixed
    iparm[px.IPARM_VERBOSE] = px.PastixVerboseNot
    pd = px.init(iparm, dparm)
    px.analyze_grid(pd, spm, nx, nx, nx, dof=dof)
    for k, v in (opts or {}).items():
        px.set_option(pd, k, float(v))
    px.set_option(pd, "model_log", 1)
    px.task_numfact(pd, spm)
    rec = timing_log(pd)
    px.finalize(pd)
    return rec[rec["time"] > 0]


# Engine options under which every blok update runs on one kernel class
# (where it fits): the built-in cut-offs, all tall tiles, all 64 x 64 tiles,
# the one-wavefront kernel over its whole range.
REGIMES = {"cut-offs": {}, "tall": {"upd_small": 0, "upd_big_rows": 1},
           "tiles64": {"upd_small": 0, "upd_big_rows": 2 ** 30},
           "wavefront": {"upd_small": 1, "upd_small_m": 256, "upd_small_n": 64}}


def class_logs(px, runs, mixed, log=print):
    """Timing-log records of the problems runs ((label, generator, grid edge,
    dof, factorization) tuples) under each regime of REGIMES, so that the
    update shapes are timed on every kernel class."""
    out = []
    for label, kind, nx, dof, facto in runs:
        for name, opts in REGIMES.items():
            rec = _logged_factorization(px, kind, nx, dof, facto, mixed, opts)
            log("logged %s (%s): %d records" % (label, name, len(rec)))
            out.append(rec)
    return np.concatenate(out)


def choice_experiment(px, train_runs, test_run, mixed=1, steps=3000, seed=7, reps=5, log=print):
    """Kernel class of every update chosen three ways for the test problem:
    the engine's cut-offs, the per-class polynomial models (the models.c
    fallback) and the learned model, both fitted to the training problems'
    logs.  Each choice is timed over reps factorizations of the test problem
    (the median of the device's own factorization time) and its factor is
    kept for the caller to check.  Also scores both models on a held-out
    fifth of the training updates."""
    arith = 0 if mixed else 1
    rec = class_logs(px, train_runs, mixed, log=log)
    upd = rec[rec["ktype"] >= KTYPE["GEMMCblk1d1d"]]
    rng = np.random.default_rng(seed)
    idx = rng.permutation(len(upd))
    ntr = int(0.8 * len(upd))
    tr, te = upd[idx[:ntr]], upd[idx[ntr:]]
    poly = class_models(tr, arith)
    learned = LearnedModel(tr, arith, steps=steps, seed=seed)
    pp = np.array([poly[int(r["variant"])].cost(arith, int(r["ktype"]), int(r["m"]), int(r["n"]), int(r["k"]))
                   for r in te])
    held_out = dict(records=int(len(te)), polynomial=error_summary(pp, te["time"]),
                    learned=error_summary(learned.predict(te), te["time"]))
    poly = class_models(upd, arith)
    learned = LearnedModel(upd, arith, steps=steps, seed=seed)
    label, kind, nx, dof, facto = test_run
    if kind == "lap7":
        spm = px.Spm.laplacian7(nx, nx, nx, 6.0, 1.0)
    elif kind == "convdiff27":
        spm = px.Spm.convdiff27(nx, nx, nx, 1.0, 0.5, 0)
    else:
        spm = px.Spm.vlaplacian7(nx, nx, nx, dof, 1.0, 0)
    iparm, dparm = px.initParam()
    iparm[px.IPARM_ORDERING] = px.PastixOrderPersonal
    iparm[px.IPARM_FACTORIZATION] = px.PastixFactGETRF if facto == "GETRF" else px.PastixFactPOTRF
    iparm[px.IPARM_MIXED] = mixed
    iparm[px.IPARM_VERBOSE] = px.PastixVerboseNot
    pd = px.init(iparm, dparm)
    px.analyze_grid(pd, spm, nx, nx, nx, dof=dof)
    px.task_numfact(pd, spm)
    out = dict(train=[r[0] for r in train_runs], train_update_records=int(len(upd)), test_problem=label,
               held_out=held_out, choices={})
    for name, kw in (("cut-offs", {}), ("polynomial", dict(poly=poly)), ("learned", dict(learned=learned))):
        pred = choose_update_kernels(pd, arith, **kw)
        times = []
        for _ in range(reps):
            px.task_numfact(pd, spm)
            times.append(px.get_stats(pd)["fact_seconds"])
        out["choices"][name] = dict(fact_seconds=float(np.median(times)), predicted_update_seconds=pred)
        log("%s: factorization %.2f ms" % (name, 1e3 * float(np.median(times))))
    return out, pd, spm, dparm


def _families(rec):
    fams = dict(panel=rec["ktype"] <= KTYPE["SYTRF"],
                trsm=(rec["ktype"] >= KTYPE["TRSMCblk1d"]) & (rec["ktype"] < KTYPE["GEMMCblk1d1d"]),
                update=rec["ktype"] >= KTYPE["GEMMCblk1d1d"])
    fams["all"] = np.ones(len(rec), dtype=bool)
    return fams


def transfer(steps=3000, lr=1e-2, seed=7, log=print):
    """Fit on the training problems' logs, score on the test problem's log;
    per factor precision (arith 0: FP32, 1: FP64) and kernel family."""
    import pastix_amd as px
    logs = {0: [], 1: []}
    runs = []
    for label, kind, nx, dof, facto, ariths in TRAIN_RUNS:
        for arith in ariths:
            rec = _logged_factorization(px, kind, nx, dof, facto, mixed=int(arith == 0))
            logs[arith].append(rec)
            runs.append(dict(problem=label, arith=arith, records=int(len(rec)),
                             logged_seconds=float(rec["time"].sum())))
            log("logged %s arith %d: %d records" % (label, arith, len(rec)))
    label, kind, nx, dof, facto, ariths = TEST_RUN
    out = dict(train=runs, test_problem=label, results={})
    cpu, gpu = Model.default_cpu(), Model.default_gpu()
    for arith in ariths:
        test = _logged_factorization(px, kind, nx, dof, facto, mixed=int(arith == 0))
        log("logged %s arith %d: %d records" % (label, arith, len(test)))
        train = np.concatenate(logs[arith])
        poly = Model().fit(train, arith, name="fit on the training problems")
        learned = LearnedModel(train, arith, steps=steps, lr=lr, seed=seed)
        preds = dict(polynomial=poly.costs(arith, test), learned=learned.predict(test),
                     default_gpu=gpu.costs(arith, test), default_cpu=cpu.costs(arith, test))
        res = dict(train_records=int(len(train)), test_records=int(len(test)),
                   test_logged_seconds=float(test["time"].sum()), mlp_train_loss=learned.loss)
        for fam, sel in _families(test).items():
            if not sel.any():
                continue
            res[fam] = {}
            for name, p in preds.items():
                # a model without coefficients for a kernel returns -1 there
                has = sel & ( p >= 0 )
                if has.any():
                    res[fam][name] = error_summary(p[has], test["time"][has])
                    res[fam][name]["coverage"] = float(has.sum() / sel.sum())
                else:
                    res[fam][name] = dict(coverage=0.0)
        out["results"]["fp32" if arith == 0 else "fp64"] = res
    return out


def main(argv=None):
    import pastix_amd as px
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--transfer", action="store_true",
                    help="fit on the 20^3 / 100^3 Laplacian and 64^3 LU logs, test on the 60^3 x 3 dof log")
    ap.add_argument("--choice", action="store_true",
                    help="fit the per-class models on the 20^3 / 100^3 Laplacian and 64^3 LU logs, then let "
                         "the cut-offs, the polynomials and the perceptron choose the update kernels of the "
                         "60^3 x 3 dof factorization, and time each choice")
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--nx", type=int, default=100)
    ap.add_argument("--mixed", type=int, default=1, help="1: FP32 factor, 0: FP64")
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--csv", default=None, help="write the timing log as model.csv here")
    ap.add_argument("--model-out", default=None, help="write the fitted polynomial model here")
    a = ap.parse_args(argv)
    if a.transfer:
        import sys
        out = transfer(steps=a.steps, log=lambda m: print(m, file=sys.stderr, flush=True))
        print(json.dumps(out, indent=1))
        return
    if a.choice:
        import sys
        train = [("20^3 Laplacian", "lap7", 20, 1, "POTRF"), ("100^3 Laplacian", "lap7", 100, 1, "POTRF"),
                 ("64^3 convection-diffusion", "convdiff27", 64, 1, "GETRF")]
        test = ("60^3 x 3 dof vector Laplacian", "vlap7", 60, 3, "POTRF")
        out, pd, spm, dparm = choice_experiment(px, train, test, mixed=a.mixed, steps=a.steps, reps=a.reps,
                                                log=lambda m: print(m, file=sys.stderr, flush=True))
        print(json.dumps(out, indent=1))
        px.finalize(pd)
        return
    spm = px.Spm.laplacian7(a.nx, a.nx, a.nx, 6.0, 1.0)
    iparm, dparm = px.initParam()
    iparm[px.IPARM_ORDERING] = px.PastixOrderPersonal
    iparm[px.IPARM_FACTORIZATION] = px.PastixFactPOTRF
    iparm[px.IPARM_MIXED] = a.mixed
    iparm[px.IPARM_VERBOSE] = px.PastixVerboseNot
    pd = px.init(iparm, dparm)
    px.analyze_grid(pd, spm, a.nx, a.nx, a.nx)
    px.set_option(pd, "model_log", 1)
    px.task_numfact(pd, spm)
    rec = timing_log(pd)
    arith = 0 if a.mixed else 1
    if a.csv:
        write_model_csv(pd, a.csv)
    out, poly, _ = compare(rec, arith, steps=a.steps)
    if a.model_out:
        Model().fit(rec, arith, name="MI355X gfx950 - pastix_amd %s" % ("fp32" if a.mixed else "fp64")) \
            .write(a.model_out)
    out.update(nx=a.nx, arith=arith, records=int(len(rec)), logged_seconds=float(rec["time"].sum()))
    print(json.dumps(out, indent=1))
    px.finalize(pd)


if __name__ == "__main__":
    main()
