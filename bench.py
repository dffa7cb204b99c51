"""Benchmark: time-to-solution of the sopalin step on the 100^3 7-point
Laplacian (4 RHS), FP32 factorization + FP64 GMRES refinement vs plain FP64,
stopping at backward error ||b-Ax||_inf/(||A||_inf||x||_inf+||b||_inf) <= 1e-12.

One step = numerical factorization (fill + factor on the GPU) + solve +
refinement of all right-hand sides, with the matrix, the structure and the
right-hand sides already resident in HBM.  The analysis (grid nested
dissection, symbolic factorization, reordering, split) is done once before
timing, as the reference's ordering_grid example does.

The line also carries the supernodal-update rooflines of both paths
(matrix-core peak), the HBM rates of one solve and one SpMV, the forward
error of the FP32+IR solution against the FP64 factorization's, and the
CPU baseline: the oracle's static-scheduler FP64 Cholesky + solve of the
same 100^3 problem on the host cores this process may use (affinity mask
capped by the cgroup quota), after a small warm-up factorization
(`--cpu-reps` timed repetitions, default 3, about 11 s each on 16 cores;
the line gives their median and minimum).

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
    ap.add_argument("--cpu-grid", type=int, default=None,
                    help="grid edge of the CPU baseline (default: the workload's own, --n)")
    ap.add_argument("--cpu-reps", type=int, default=3,
                    help="timed repetitions of the CPU baseline (after a small warm-up factorization)")
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


def host_cores():
    """Threads for the CPU baseline: the CPUs of this process's affinity mask,
    capped by the cgroup CPU quota when one is set (a GPU box grants each GPU
    a share of its host cores).  Returns (threads, affinity, quota, model)."""
    aff = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
        if q != "max":
            quota = max(1, int(int(q) // int(per)))
    except (OSError, ValueError):
        pass
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return (min(aff, quota) if quota else aff), aff, quota, model


def cpu_baseline(px, ngrid, nrhs, reps):
    """The oracle's static-scheduler FP64 Cholesky + nrhs-RHS solve of the
    ngrid^3 Laplacian on the host cores (same analysis as the GPU paths):
    one small warm-up factorization, then reps timed repetitions (min and
    median)."""
    import gc
    import oracle_ref as orc
    threads, aff, quota, model = host_cores()

    def setup(g):
        spm = px.Spm.laplacian7(g, g, g)
        iparm, dparm = px.initParam()
        iparm[px.IPARM_ORDERING] = px.PastixOrderPersonal
        iparm[px.IPARM_FACTORIZATION] = px.PastixFactPOTRF
        iparm[px.IPARM_VERBOSE] = px.PastixVerboseNot
        pd = px.init(iparm, dparm)
        px.analyze_grid(pd, spm, g, g, g)
        px.lib.pastix_subtask_spm2bcsc(pd, ctypes.byref(spm.s))
        tau = np.finfo(np.float64).eps / 2 * dparm[px.DPARM_A_NORM]
        return spm, pd, dparm, tau

    spm, pd, dparm, tau = setup(min(ngrid, 30))
    orc.OracleFactor(px.solver_export(pd), px.bcsc_export(pd), tau, np.float64, nthreads=threads)
    px.finalize(pd)
    spm, pd, dparm, tau = setup(ngrid)
    s = px.solver_export(pd)
    bc = px.bcsc_export(pd)
    fl = dparm[px.DPARM_FACT_THFLOPS]
    _, b = spm.gen_rhs(nrhs, seed=0)
    perm = px.order_export(pd)[0]
    bp = np.zeros_like(b)
    bp[perm] = b
    px.finalize(pd)
    tf, ts = [], []
    for r in range(max(1, reps)):
        F = orc.OracleFactor(s, bc, tau, np.float64, nthreads=threads)
        t0 = time.perf_counter()
        F.solve(bp)
        ts.append(time.perf_counter() - t0)
        tf.append(F.seconds)
        log("cpu rep %d: factorization %.2f s, solve %.2f s" % (r, tf[-1], ts[-1]))
        del F
        gc.collect()
    tts = [a + c for a, c in zip(tf, ts)]
    med = float(np.median(tts))
    return dict(value=fl / float(np.median(tf)) / 1e9, unit="GFLOP/s", cores=threads, kind="port",
                sample="FP64 Cholesky + %d-RHS solve of the %d^3 7-point Laplacian (%d unknowns, %.3e flop), "
                       "the oracle's static scheduler with %s on %d threads (%s; affinity %d CPUs, cgroup quota "
                       "%s); warm-up then %d repetition(s): factorization median %.2f s min %.2f s, solve median "
                       "%.2f s (the oracle's sequential sweeps), time-to-solution median %.2f s min %.2f s"
                       % (nrhs, ngrid, ngrid ** 3, fl, orc.blas_name(), threads, model, aff,
                          quota if quota else "none", len(tf), float(np.median(tf)), min(tf),
                          float(np.median(ts)), med, min(tts)),
                fact_seconds=float(np.median(tf)), solve_seconds=float(np.median(ts)), tts_seconds=med,
                tts_min_seconds=min(tts), thflops=fl, reps=len(tf))


def solve_and_spmv_rates(px, run, reps=5):
    """Device time of one solve (HIP events inside the engine) and of one
    SpMV of all right-hand sides, with their algorithmic HBM bytes.

    Solve bytes: each sweep reads every off-diagonal tile of the factor once
    (sum over cblks of n_k (stride_k - n_k) entries) and the lower triangle of
    each diagonal block's inverse (n_k (n_k + 1) / 2 entries), in the factor
    precision; the vectors (n x nrhs, read and written a few times) are below
    1 % of it and left out.  SpMV bytes: the stored CSC once (8-byte values
    and row indices, n + 1 column pointers) plus x read and y written once."""
    s = px.solver_export(run.pd)
    cb = s["cblk"].reshape(-1, 6)[:s["cblknbr"]]
    w = cb[:, 1] - cb[:, 0] + 1
    offd = float(np.sum(w * (cb[:, 3] - w)))
    tri = float(np.sum(w * (w + 1) / 2))
    es = 4 if run.iparm[px.IPARM_MIXED] else 8
    solve_bytes = 2.0 * es * (offd + tri)
    times = []
    for _ in range(reps):
        px.lib.pxd_vec_copy(run.dev, run.N, run.nrhs, run.db, run.dx)
        px._check("solve", px.lib.pastix_amd_solve_dev(run.pd, run.nrhs, run.dx, run.N))
        times.append(px.get_stats(run.pd)["solve_seconds"])
    t_solve = float(np.median(times))
    dy = ctypes.c_void_p()
    px.lib.pxd_malloc(run.dev, run.nbytes, ctypes.byref(dy))
    nsp = 20
    px.lib.pxd_sync(run.dev)
    t0 = time.perf_counter()
    for _ in range(nsp):
        px.lib.pxd_spmv(run.dev, 1.0, run.dx, 0.0, dy, run.nrhs, run.N)
    px.lib.pxd_sync(run.dev)
    t_spmv = (time.perf_counter() - t0) / nsp
    px.lib.pxd_free(run.dev, dy)
    bc_n, bc_nnz = px.pastix_int(), px.pastix_int()
    px.lib.pastix_amd_bcsc_sizes(run.pd, ctypes.byref(bc_n), ctypes.byref(bc_nnz))
    spmv_bytes = bc_nnz.value * 16.0 + (run.N + 1) * 8.0 + 2.0 * run.N * run.nrhs * 8.0
    return {
        "solve": {"bound": "hbm", "achieved": solve_bytes / t_solve / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": solve_bytes / t_solve / 1e9 / HBM_PEAK_GBS, "seconds": t_solve,
                  "algorithmic_bytes": solve_bytes,
                  "kernel": "forward + backward sweeps of one %d-RHS solve (event-timed)" % run.nrhs},
        "spmv": {"bound": "hbm", "achieved": spmv_bytes / t_spmv / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": spmv_bytes / t_spmv / 1e9 / HBM_PEAK_GBS, "seconds": t_spmv,
                 "algorithmic_bytes": spmv_bytes,
                 "kernel": "k_spmv<double, double>, %d right-hand sides (mean of %d launches)" % (run.nrhs, nsp)},
    }


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

    results, xs = {}, {}
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
        # per-phase timing pass (one stream, events around each phase)
        px.set_option(run.pd, "timing", 1)
        run.step()
        st = px.get_stats(run.pd)
        px.set_option(run.pd, "timing", 0)
        r["kernel_seconds"] = {k: st[k] for k in ("fill_seconds", "panel_seconds", "trsm_seconds",
                                                  "update_seconds")}
        r["update_flops"] = update_flops(px, run.pd)
        r["bandwidth"] = solve_and_spmv_rates(px, run)
        xs[label] = r.pop("x")
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
        g = args.cpu_grid if args.cpu_grid else n
        log("step: CPU baseline (%d^3, %d repetition(s))" % (g, args.cpu_reps))
        cpu = cpu_baseline(px, g, args.nrhs, args.cpu_reps)
        log("cpu: %.1f GFLOP/s on %d threads, time-to-solution %.2f s" % (cpu["value"], cpu["cores"],
                                                                        cpu["tts_seconds"]))

    head = results["fp32_ir"]
    fl = head["thflops"]

    def update_roofline(r, prec):
        upd = r.get("update_flops", 0.0)
        upd_t = r.get("kernel_seconds", {}).get("update_seconds", 0.0)
        ach = upd / upd_t / 1e12 if upd_t > 0 else 0.0
        return ach, upd

    achieved, upd = update_roofline(head, "f32")
    traffic = traffic64 = traffic_launch = None
    if args.traffic and os.path.exists(args.traffic):
        with open(args.traffic) as f:
            tj = json.load(f)
        traffic = tj.get("hbm_bytes_per_factorization")
        traffic64 = tj.get("fp64_hbm_bytes_per_factorization")
        if tj.get("fetch_bytes_per_launch_k_update_dma"):
            traffic_launch = tj["fetch_bytes_per_launch_k_update_dma"] + tj["write_bytes_per_launch_k_update_dma"]
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
                     "traffic_scope": "HBM bytes of the update launches of one factorization (rocprofv3 FETCH_SIZE "
                                      "+ WRITE_SIZE, copy-calibrated)", "traffic_per_launch": traffic_launch,
                     "kernel": "k_update_dma<8,3> + k_update_mfma<float> + k_update_small<float> (supernodal "
                               "update, all launches of one factorization)", "algorithmic_flops": upd},
        "cpu_baseline": None,
    }
    out["roofline_solve_fp32"] = head["bandwidth"]["solve"]
    out["roofline_spmv"] = head["bandwidth"]["spmv"]
    if "fp64" in results:
        r64 = results["fp64"]
        out["fp64"] = r64
        out["speedup_fp32ir_vs_fp64"] = r64["tts"] / head["tts"]
        a64, u64 = update_roofline(r64, "f64")
        out["roofline_fp64"] = {"bound": "mfma", "achieved": a64, "peak": MFMA_PEAK_TFLOPS["f64"], "unit": "TFLOP/s",
                                "frac": a64 / MFMA_PEAK_TFLOPS["f64"], "traffic": traffic64,
                                "kernel": "k_update_big<double> + k_update_mfma<double> + k_update_small<double> "
                                          "(all launches of one factorization)", "algorithmic_flops": u64}
        out["roofline_solve_fp64"] = r64["bandwidth"]["solve"]
        # forward error of the FP32+IR solution against the FP64 factorization's
        x32, x64 = xs["fp32_ir"], xs["fp64"]
        out["ferr_vs_fp64"] = float(max(np.abs(x32[:, c] - x64[:, c]).max() / np.abs(x64[:, c]).max()
                                        for c in range(x64.shape[1])))
    if cpu is not None:
        out["cpu_baseline"] = {k: cpu[k] for k in ("value", "unit", "cores", "kind", "sample")}
        if cpu["reps"] and (args.cpu_grid in (None, n)):
            out["cpu_tts_seconds"] = cpu["tts_seconds"]
            out["speedup_vs_cpu"] = cpu["tts_seconds"] / head["tts"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
