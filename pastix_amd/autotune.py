"""Autotuner of the engine's schedule options on one MI355X.

It measures the time-to-solution (numerical factorization + solve +
refinement, matrix and right-hand sides resident in HBM) of a set of
problems under candidate values of the engine options that choose tile
classes, cut-offs and launch shapes without changing the results beyond
rounding:

- ``upd_big_rows``: stacked rows from which a blok update runs on the tall
  256/128-row tiles instead of 64 x 64 tiles;
- ``upd_small``, ``upd_small_m``, ``upd_small_n``: the one-wavefront kernel
  for updates of at most upd_small_m rows, upd_small_n columns and 32 k
  values;
- ``solve_wide_max``: level-solve launches from which the tiles load 16
  columns per batch instead of all of them;
- ``panel_mw``: the multi-workgroup Cholesky panel;
- ``upd_dma``: single-precision tall tiles on the kernel that loads its
  operands straight into LDS;
- ``upd_dma_pad``: extra LDS per workgroup of that kernel, which sets how
  many of its workgroups share a CU with the lookahead stream's panels;
- ``solve_split_max``, ``solve_split_wg``: how many workgroups share a tile
  of a level-solve launch with few tiles;
- ``solve_spread``: the fewest tiles of a cblk whose backward gather goes
  through partial-sum copies (0: none).

The search is coordinate descent from the current values (each option in
turn over its candidates, the others fixed, two passes); the objective is
the sum over the problems of the median time-to-solution of ``--steps``
steps.  The chosen values go to ``tuning_mi355x.txt`` next to the library,
which every engine context reads when it is created (``pxd_create``), with
the measurements as comments.

    python -m pastix_amd.autotune [--problems c2,c3] [--steps 3] [--out FILE]
"""
import argparse
import ctypes
import os
import sys
import time

import numpy as np

from . import (PastixError, Spm, analyze_grid, init, initParam, finalize, lib, _check, set_option,
               order_export)
from . import (IPARM_ORDERING, IPARM_FACTORIZATION, IPARM_MIXED, IPARM_VERBOSE, IPARM_REFINEMENT,
               PastixOrderPersonal, PastixFactPOTRF, PastixFactGETRF, PastixVerboseNot, PastixRefineGMRES)

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "tuning_mi355x.txt")

SPACE = {
    "upd_big_rows": [48, 64, 96, 128, 192, 256, 384],
    "upd_small": [0, 1],
    "solve_wide_max": [32, 64, 128, 256, 1024],
    "panel_mw": [0, 1],
    "upd_dma": [0, 1],
    "upd_dma_pad": [0, 2048, 4096, 6144, 8192],
    "upd_small_m": [64, 128, 256],
    "upd_small_n": [16, 32, 64],
    "solve_split_max": [1, 2, 4, 8],
    "solve_split_wg": [256, 512, 1024, 2048],
    "solve_spread": [0, 8, 16, 32],
}
DEFAULTS = {"upd_big_rows": 192, "upd_small": 1, "solve_wide_max": 128, "panel_mw": 1, "upd_dma": 1,
            "upd_dma_pad": 4096, "upd_small_m": 64, "upd_small_n": 16, "solve_split_max": 4,
            "solve_split_wg": 512, "solve_spread": 16}

# problems: (label, generator, grid edge, factorization, mixed precision, RHS)
PROBLEMS = {
    "c2": ("100^3 Laplacian, FP32 Cholesky + GMRES", "lap7", 100, "potrf", 1, 4),
    "c2d": ("100^3 Laplacian, FP64 Cholesky", "lap7", 100, "potrf", 0, 4),
    "c3": ("64^3 convection-diffusion, FP32 LU + GMRES", "convdiff27", 64, "getrf", 1, 8),
    "small": ("40^3 Laplacian, FP32 Cholesky + GMRES", "lap7", 40, "potrf", 1, 4),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class Problem:
    """One analysed problem with b and x resident on the device."""

    def __init__(self, key):
        desc, gen, nx, facto, mixed, nrhs = PROBLEMS[key]
        self.key, self.desc = key, desc
        if gen == "lap7":
            self.spm = Spm.laplacian7(nx, nx, nx)
        else:
            self.spm = Spm.convdiff27(nx, nx, nx, 1.0, 0.5, 0)
        iparm, dparm = initParam()
        iparm[IPARM_ORDERING] = PastixOrderPersonal
        iparm[IPARM_FACTORIZATION] = PastixFactGETRF if facto == "getrf" else PastixFactPOTRF
        iparm[IPARM_MIXED] = mixed
        iparm[IPARM_VERBOSE] = PastixVerboseNot
        iparm[IPARM_REFINEMENT] = PastixRefineGMRES
        self.iparm, self.dparm = iparm, dparm
        self.pd = init(iparm, dparm)
        analyze_grid(self.pd, self.spm, nx, nx, nx)
        rc = lib.pastix_subtask_spm2bcsc(self.pd, ctypes.byref(self.spm.s))
        rc |= lib.pastix_subtask_bcsc2ctab(self.pd)
        if rc:
            raise PastixError("bcsc upload", rc)
        set_option(self.pd, "berr_target", 1e-12)
        self.n, self.nrhs = self.spm.n, nrhs
        _, b = self.spm.gen_rhs(nrhs, seed=0)
        perm = order_export(self.pd)[0]
        bp = np.zeros_like(b)
        bp[perm] = b
        bp = np.asfortranarray(bp)
        self.dev = ctypes.c_void_p()
        _check("device", lib.pastix_amd_device_handle(self.pd, ctypes.byref(self.dev)))
        self.db, self.dx = ctypes.c_void_p(), ctypes.c_void_p()
        lib.pxd_malloc(self.dev, bp.nbytes, ctypes.byref(self.db))
        lib.pxd_malloc(self.dev, bp.nbytes, ctypes.byref(self.dx))
        lib.pxd_h2d(self.dev, self.db, bp.ctypes.data_as(ctypes.c_void_p), bp.nbytes)

    def step(self):
        t0 = time.perf_counter()
        _check("sopalin", lib.pastix_subtask_sopalin(self.pd))
        lib.pxd_vec_copy(self.dev, self.n, self.nrhs, self.db, self.dx)
        _check("solve", lib.pastix_amd_solve_dev(self.pd, self.nrhs, self.dx, self.n))
        it = ctypes.c_int64()
        _check("refine", lib.pastix_amd_refine_dev(self.pd, self.nrhs, self.db, self.dx, ctypes.byref(it)))
        lib.pxd_sync(self.dev)
        return time.perf_counter() - t0

    def measure(self, opts, steps):
        for k, v in opts.items():
            set_option(self.pd, k, float(v))
        self.step()   # warm-up (and the task lists of new options)
        return float(np.median([self.step() for _ in range(steps)]))

    def close(self):
        lib.pxd_free(self.dev, self.db)
        lib.pxd_free(self.dev, self.dx)
        finalize(self.pd)


def tune(keys, steps, passes=2):
    probs = [Problem(k) for k in keys]
    cur = dict(DEFAULTS)
    cache = {}

    def cost(opts):
        key = tuple(sorted(opts.items()))
        if key not in cache:
            per = [p.measure(opts, steps) for p in probs]
            cache[key] = (sum(per), per)
            log("  %s -> %.2f ms (%s)" % (dict(opts), 1e3 * sum(per), ", ".join("%.2f" % (1e3 * t) for t in per)))
        return cache[key]

    base = cost(cur)
    best = base
    for p in range(passes):
        changed = False
        for name, values in SPACE.items():
            for v in values:
                if v == cur[name]:
                    continue
                trial = dict(cur)
                trial[name] = v
                c = cost(trial)
                if c[0] < best[0] * 0.995:   # a change must gain more than the noise
                    best, cur, changed = c, trial, True
        if not changed:
            break
    # the defaults again, after the search (the same clock conditions)
    cache.pop(tuple(sorted(DEFAULTS.items())), None)
    base = cost(dict(DEFAULTS))
    for p in probs:
        p.close()
    return cur, best, base


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--problems", default="c2,c2d,c3")
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--out", default=DEFAULT_OUT)
    args = ap.parse_args()
    keys = args.problems.split(",")
    # measure from the built-in defaults, not from a previous tuning file
    os.environ["PXD_TUNING_FILE"] = os.devnull
    cur, best, base = tune(keys, args.steps)
    lines = ["# Engine options chosen by pastix_amd/autotune.py on one MI355X (read by pxd_create).",
             "# Objective: sum of the median time-to-solution over: %s" % "; ".join(PROBLEMS[k][0] for k in keys),
             "# built-in defaults %s: %.2f ms (%s)" % (DEFAULTS, 1e3 * base[0],
                                                        ", ".join("%.2f" % (1e3 * t) for t in base[1])),
             "# chosen: %.2f ms (%s)" % (1e3 * best[0], ", ".join("%.2f" % (1e3 * t) for t in best[1]))]
    lines += ["%s %g" % (k, v) for k, v in sorted(cur.items())]
    with open(args.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


if __name__ == "__main__":
    main()
