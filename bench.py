"""Benchmark: time-to-solution of the sopalin step on the 100^3 7-point
Laplacian (4 RHS), FP32 factorization + FP64 GMRES refinement vs plain FP64,
stopping at backward error ||b-Ax||_inf/(||A||_inf||x||_inf+||b||_inf) <= 1e-12.

One step = numerical factorization (fill + factor on the GPU) + solve +
refinement of all right-hand sides, with the matrix, the structure and the
right-hand sides already resident in HBM.  The analysis (grid nested
dissection, symbolic factorization, reordering, split) is done once before
timing, as the reference's ordering_grid example does.

Prints one JSON line (the driver's contract).  `--gpus` above 1 is refused:
the engine drives one MI355X.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0
MFMA_PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=100, help="grid edge (100 -> 1,000,000 unknowns)")
    ap.add_argument("--nrhs", type=int, default=4)
    ap.add_argument("--berr", type=float, default=1e-12)
    ap.add_argument("--cpu-sample", type=int, default=60, help="grid edge of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS),
                    help="c2 (default): the headline 100^3 Cholesky line; c1, c3, c4, c5: the other "
                         "configurations, each on its own JSON line")
    ap.add_argument("--opt", action="append", default=[], metavar="KEY=VALUE",
                    help="engine option for both paths (e.g. prio=0, upd_big_rows=256)")
    ap.add_argument("--profile-stats", action="store_true", help="per-kernel-class event timing pass")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="after the timed region, one device copy of this many MiB (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters of a profiled run)")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "update_traffic_100cubed.json"),
                    help="measured HBM bytes of the update kernels (from the counter runs), if present")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class Run:
    """One analysed problem with the matrix and the RHS resident on the device."""

    def __init__(self, px, spm, n, nrhs, mixed, berr, opts=(), facto=None, refine=None, dof=1):
        self.px = px
        self.spm = spm
        iparm, dparm = px.initParam()
        iparm[px.IPARM_ORDERING] = px.PastixOrderPersonal
        iparm[px.IPARM_FACTORIZATION] = px.PastixFactPOTRF if facto is None else facto
        if refine is not None:
            iparm[px.IPARM_REFINEMENT] = refine
        iparm[px.IPARM_MIXED] = mixed
        iparm[px.IPARM_VERBOSE] = px.PastixVerboseNot
        self.iparm, self.dparm = iparm, dparm
        self.pd = px.init(iparm, dparm)
        t0 = time.perf_counter()
        px.analyze_grid(self.pd, spm, n, n, n, dof=dof)
        self.t_analyze = time.perf_counter() - t0
        rc = px.lib.pastix_subtask_spm2bcsc(self.pd, ctypes.byref(spm.s))
        rc |= px.lib.pastix_subtask_bcsc2ctab(self.pd)
        if rc:
            raise RuntimeError("bcsc upload failed (%d)" % rc)
        px.set_option(self.pd, "berr_target", berr)
        for kv in opts:
            k, v = kv.split("=")
            px.set_option(self.pd, k, float(v))
        self.nrhs = nrhs
        self.N = spm.n
        self.x_true, b = spm.gen_rhs(nrhs, seed=0)
        perm = px.order_export(self.pd)[0]
        self.perm = perm
        bp = np.zeros_like(b)
        bp[perm] = b
        self.b = b
        dev = ctypes.c_void_p()
        px._check("device", px.lib.pastix_amd_device_handle(self.pd, ctypes.byref(dev)))
        self.dev = dev
        nbytes = bp.nbytes
        self.db = ctypes.c_void_p()
        self.dx = ctypes.c_void_p()
        px.lib.pxd_malloc(dev, nbytes, ctypes.byref(self.db))
        px.lib.pxd_malloc(dev, nbytes, ctypes.byref(self.dx))
        bp = np.asfortranarray(bp)
        px.lib.pxd_h2d(dev, self.db, bp.ctypes.data_as(ctypes.c_void_p), nbytes)
        self.nbytes = nbytes

    def step(self):
        """factorize + solve + refine; returns (t_fact, t_solve, t_refine, iters)."""
        px = self.px
        t0 = time.perf_counter()
        px._check("sopalin", px.lib.pastix_subtask_sopalin(self.pd))
        t1 = time.perf_counter()
        px.lib.pxd_vec_copy(self.dev, self.N, self.nrhs, self.db, self.dx)
        px._check("solve", px.lib.pastix_amd_solve_dev(self.pd, self.nrhs, self.dx, self.N))
        px.lib.pxd_sync(self.dev)
        t2 = time.perf_counter()
        it = px.pastix_int()
        px._check("refine", px.lib.pastix_amd_refine_dev(self.pd, self.nrhs, self.db, self.dx, ctypes.byref(it)))
        px.lib.pxd_sync(self.dev)
        t3 = time.perf_counter()
        return t1 - t0, t2 - t1, t3 - t2, it.value

    def solution(self):
        xp = np.zeros((self.N, self.nrhs), dtype=self.spm.dtype, order="F")
        self.px.lib.pxd_d2h(self.dev, xp.ctypes.data_as(ctypes.c_void_p), self.dx, self.nbytes)
        return xp[self.perm]

    def close(self):
        self.px.lib.pxd_free(self.dev, self.db)
        self.px.lib.pxd_free(self.dev, self.dx)
        self.px.finalize(self.pd)


def update_flops(px, pd):
    """Algorithmic flops of all supernodal updates: sum 2*M_i*N_b*K."""
    s = px.solver_export(pd)
    cb = s["cblk"].reshape(-1, 6)
    bk = s["blok"].reshape(-1, 5)
    tot = 0.0
    for k in range(s["cblknbr"]):
        K = cb[k, 1] - cb[k, 0] + 1
        rows = bk[cb[k, 2]:cb[k + 1, 2], 1] - bk[cb[k, 2]:cb[k + 1, 2], 0] + 1
        below = np.cumsum(rows[::-1])[::-1]      # rows from blok b to the end
        tot += float(np.sum(2.0 * rows[1:] * below[1:] * K))
    return tot


def measure(run, steps, warmup):
    for _ in range(warmup):
        run.step()
    rec = []
    for _ in range(steps):
        rec.append(run.step())
    x = run.solution()
    berr, each = run.spm.backward_error(run.b, x)
    ferr = float(np.abs(x - run.x_true).max() / np.abs(run.x_true).max())
    a = np.array([r[:3] for r in rec])
    return dict(fact=float(np.median(a[:, 0])), solve=float(np.median(a[:, 1])),
                refine=float(np.median(a[:, 2])), tts=float(np.median(a.sum(axis=1))),
                tts_min=float(a.sum(axis=1).min()), iters=int(rec[-1][3]),
                berr=float(berr), ferr=ferr, x=x)


def cpu_baseline(px, nsample, nrhs):
    """The oracle's static-scheduler Cholesky + solve on the host cores, on a
    bounded sample of the workload (a nsample^3 Laplacian, same pipeline)."""
    import oracle_ref as orc
    spm = px.Spm.laplacian7(nsample, nsample, nsample)
    iparm, dparm = px.initParam()
    iparm[px.IPARM_ORDERING] = px.PastixOrderPersonal
    iparm[px.IPARM_FACTORIZATION] = px.PastixFactPOTRF
    pd = px.init(iparm, dparm)
    px.analyze_grid(pd, spm, nsample, nsample, nsample)
    px.lib.pastix_subtask_spm2bcsc(pd, ctypes.byref(spm.s))
    s = px.solver_export(pd)
    bc = px.bcsc_export(pd)
    cores = len(os.sched_getaffinity(0))
    cores = min(cores, 16)
    tau = np.finfo(np.float64).eps / 2 * dparm[px.DPARM_A_NORM]
    F = orc.OracleFactor(s, bc, tau, np.float64, nthreads=cores)
    fl = dparm[px.DPARM_FACT_THFLOPS]
    _, b = spm.gen_rhs(nrhs, seed=0)
    perm = px.order_export(pd)[0]
    bp = np.zeros_like(b)
    bp[perm] = b
    t0 = time.perf_counter()
    F.solve(bp)
    tsolve = time.perf_counter() - t0
    px.finalize(pd)
    return dict(value=fl / F.seconds / 1e9, unit="GFLOP/s", cores=cores, kind="port",
                sample="FP64 Cholesky of the %d^3 7-point Laplacian (%d unknowns, %.3e flop) with the "
                       "oracle's static scheduler (%s): factorization %.2f s, %d-RHS solve %.2f s"
                       % (nsample, nsample ** 3, fl, orc.blas_name(), F.seconds, nrhs, tsolve),
                fact_seconds=F.seconds, solve_seconds=tsolve, thflops=fl)


# The other configurations of the benchmark suite (the headline is c2):
# generator, grid edge, unknowns per node, factorization, refinement, RHS,
# backward-error target, paths (label, IPARM_MIXED).
CONFIGS = {
    "c1": dict(desc="20^3 7-point Laplacian, FP64 Cholesky, 1 RHS", gen=("lap7",), nx=20, dof=1,
               facto="PastixFactPOTRF", refine=None, nrhs=1, berr=1e-13, paths=(("fp64", 0),)),
    "c2": dict(desc="100^3 7-point Laplacian, POTRF, 4 RHS, grid nested dissection", gen=("lap7",), nx=100,
               dof=1, facto="PastixFactPOTRF", refine=None, nrhs=4, berr=1e-12,
               paths=(("fp32_ir", 1), ("fp64", 0))),
    "c3": dict(desc="64^3 27-point convection-diffusion, LU with static pivoting, FP32 LU + FP64 GMRES vs FP64, "
                    "8 RHS", gen=("convdiff27",), nx=64, dof=1, facto="PastixFactGETRF", refine="PastixRefineGMRES",
               nrhs=8, berr=1e-12, paths=(("fp32_ir", 1), ("fp64", 0))),
    "c4": dict(desc="80^3 Helmholtz A = L7 - (0.5 + 0.1i) I, complex-symmetric LDL^T, complex64 + complex128 "
                    "BiCGStab vs complex128, 16 RHS", gen=("helmholtz7",), nx=80, dof=1,
               facto="PastixFactLDLT", refine="PastixRefineBiCGSTAB", nrhs=16, berr=1e-12,
               paths=(("c64_ir", 1), ("c128", 0))),
    "c5": dict(desc="60^3 nodes x 3 dof vector Laplacian (lognormal coefficients), FP32 Cholesky + FP64 GMRES vs "
                    "FP64, 32 RHS", gen=("vlap7",), nx=60, dof=3, facto="PastixFactPOTRF",
               refine="PastixRefineGMRES", nrhs=32, berr=1e-12, paths=(("fp32_ir", 1), ("fp64", 0))),
}


def make_config_matrix(px, cfg):
    nx = cfg["nx"]
    kind = cfg["gen"][0]
    if kind == "lap7":
        return px.Spm.laplacian7(nx, nx, nx)
    if kind == "convdiff27":
        return px.Spm.convdiff27(nx, nx, nx, 1.0, 0.5, 0)
    if kind == "helmholtz7":
        return px.Spm.helmholtz7(nx, nx, nx, 0.5, 0.1)
    return px.Spm.vlaplacian7(nx, nx, nx, cfg["dof"], 1.0, 0)


def run_other_config(px, args):
    """One JSON line for a configuration other than the headline one."""
    cfg = CONFIGS[args.config]
    log("step: generate %s" % cfg["desc"])
    spm = make_config_matrix(px, cfg)
    facto = getattr(px, cfg["facto"])
    refine = getattr(px, cfg["refine"]) if cfg["refine"] else None
    results = {}
    for label, mixed in cfg["paths"]:
        log("step: analyse + upload (%s)" % label)
        run = Run(px, spm, cfg["nx"], cfg["nrhs"], mixed, cfg["berr"], args.opt, facto=facto, refine=refine,
                  dof=cfg["dof"])
        r = measure(run, args.steps, args.warmup)
        r["thflops"] = run.dparm[px.DPARM_FACT_THFLOPS]
        r["t_analyze"] = run.t_analyze
        r["static_pivots"] = int(run.iparm[px.IPARM_STATIC_PIVOTING])
        del r["x"]
        results[label] = r
        log("%s: fact %.3f s solve %.3f s refine %.3f s (%d it) tts %.3f s berr %.2e"
            % (label, r["fact"], r["solve"], r["refine"], r["iters"], r["tts"], r["berr"]))
        run.close()
    head = results[cfg["paths"][0][0]]
    out = {"metric": "fact+solve s at bwd err<=%g: %s" % (cfg["berr"], cfg["desc"]), "value": head["tts"],
           "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": head["tts"] * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
           "data": "synthetic (%s)" % cfg["gen"][0],
           "config": {"workload": args.config, "unknowns": int(spm.n), "nrhs": cfg["nrhs"],
                      "berr_target": cfg["berr"]},
           "gflops_fact": head["thflops"] / head["fact"] / 1e9}
    out.update(results)
    if len(cfg["paths"]) == 2:
        out["speedup_mixed_vs_full"] = results[cfg["paths"][1][0]]["tts"] / head["tts"]
    ok = all(r["berr"] <= cfg["berr"] for r in results.values())
    out["berr_met"] = bool(ok)
    print(json.dumps(out))
    return 0 if ok else 1


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this engine drives exactly one MI355X; --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    import pastix_amd as px
    if px.device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(3)
    if args.config != "c2":
        sys.exit(run_other_config(px, args))
    n = args.n
    log("step: generate %d^3 Laplacian" % n)
    spm = px.Spm.laplacian7(n, n, n)

    results = {}
    for label, mixed in (("fp32_ir", 1), ("fp64", 0)):
        if label == "fp64" and args.no_fp64:
            continue
        log("step: analyse + upload (%s)" % label)
        run = Run(px, spm, n, args.nrhs, mixed, args.berr, args.opt)
        log("step: analysis %.2f s; timing %d steps (%s)" % (run.t_analyze, args.steps, label))
        r = measure(run, args.steps, args.warmup)
        r["thflops"] = run.dparm[px.DPARM_FACT_THFLOPS]
        r["t_analyze"] = run.t_analyze
        r["static_pivots"] = int(run.iparm[px.IPARM_STATIC_PIVOTING])
        st = px.get_stats(run.pd)
        r["levels"] = int(st["levels"])
        r["launches"] = int(st["launches"])
        if args.profile_stats or label == "fp32_ir":
            px.set_option(run.pd, "timing", 1)
            run.step()
            st = px.get_stats(run.pd)
            px.set_option(run.pd, "timing", 0)
            r["kernel_seconds"] = {k: st[k] for k in ("fill_seconds", "panel_seconds", "trsm_seconds",
                                                      "update_seconds")}
            r["update_flops"] = update_flops(px, run.pd)
        del r["x"]
        results[label] = r
        log("%s: fact %.3f s solve %.3f s refine %.3f s (%d it) tts %.3f s berr %.2e"
            % (label, r["fact"], r["solve"], r["refine"], r["iters"], r["tts"], r["berr"]))
        run.close()

    if args.copy_probe > 0:
        # one plain device copy of known size, outside every timed region
        import ctypes as ct
        nbytes = args.copy_probe << 20
        dev = ct.c_void_p()
        iparm, dparm = px.initParam()
        pdp = px.init(iparm, dparm)
        px._check("device", px.lib.pastix_amd_device_handle(pdp, ct.byref(dev)))
        a, b = ct.c_void_p(), ct.c_void_p()
        px.lib.pxd_malloc(dev, nbytes, ct.byref(a))
        px.lib.pxd_malloc(dev, nbytes, ct.byref(b))
        px.lib.pxd_memset0(dev, a, nbytes)
        px.lib.pxd_vec_copy(dev, nbytes // 8, 1, a, b)
        px.lib.pxd_sync(dev)
        px.lib.pxd_free(dev, a)
        px.lib.pxd_free(dev, b)
        px.finalize(pdp)
        log("copy probe: %d MiB device copy done" % args.copy_probe)

    cpu = None
    if not args.no_cpu:
        log("step: CPU baseline sample")
        cpu = cpu_baseline(px, args.cpu_sample, args.nrhs)
        log("cpu: %.1f GFLOP/s on %d cores" % (cpu["value"], cpu["cores"]))

    head = results["fp32_ir"]
    fl = head["thflops"]
    upd = head.get("update_flops", 0.0)
    upd_t = head.get("kernel_seconds", {}).get("update_seconds", 0.0)
    achieved = upd / upd_t / 1e12 if upd_t > 0 else 0.0
    traffic = None
    if args.traffic and os.path.exists(args.traffic):
        with open(args.traffic) as f:
            traffic = json.load(f).get("hbm_bytes_per_factorization")
    out = {
        "metric": "fact+solve s & GFLOP/s at bwd err<=1e-12: FP32+IR vs FP64, 100^3 Laplacian",
        "value": head["tts"],
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["tts"] * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 factor + f64 refinement",
        "data": "synthetic (7-point Laplacian, x_true uniform[-1,1] from plrnt seed 0..%d)" % (args.nrhs - 1),
        "config": {"workload": "%d^3 7-point Laplacian, POTRF, %d RHS, grid nested dissection"
                               % (n, args.nrhs), "unknowns": n ** 3, "nrhs": args.nrhs,
                   "berr_target": args.berr},
        "fp32_ir": {k: v for k, v in head.items()},
        "gflops_fact": fl / head["fact"] / 1e9,
        "gflops_fact_binary": fl / head["fact"] / 1024 ** 3,
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": MFMA_PEAK_TFLOPS["f32"],
                     "unit": "TFLOP/s", "frac": achieved / MFMA_PEAK_TFLOPS["f32"], "traffic": traffic,
                     "kernel": "k_update_big<float> + k_update_mfma<float> (supernodal update, all launches "
                               "of one factorization)", "algorithmic_flops": upd},
        "cpu_baseline": None,
    }
    if "fp64" in results:
        r64 = results["fp64"]
        out["fp64"] = r64
        out["speedup_fp32ir_vs_fp64"] = r64["tts"] / head["tts"]
    if cpu is not None:
        out["cpu_baseline"] = {k: cpu[k] for k in ("value", "unit", "cores", "kind", "sample")}
        cpu_tts = fl / (cpu["value"] * 1e9) + cpu["solve_seconds"] * (n ** 3) / (args.cpu_sample ** 3)
        out["speedup_vs_cpu_projected"] = cpu_tts / head["tts"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
