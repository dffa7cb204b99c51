"""Build the in-tree shared libraries.

- ``pastix_amd/libpastix_amd.so``: the PaStiX C API (host C, gcc) and the
  MI355X engine (HIP, hipcc for gfx950), linked into one C-ABI library.
- ``oracle/liboracle.so``: the CPU restatement used by tests and by the
  benchmark's CPU baseline (plain C + pthreads; never linked into the
  product).

Run ``python pastix_amd/build.py`` (or ``__graft_entry__.build()``).  Objects
go to ``_build/`` (git-ignored); the .so files stay in-tree so that they
travel with the repository snapshot.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "pastix_amd")
HOST_DIR = os.path.join(PKG, "csrc", "host")
DEV_DIR = os.path.join(PKG, "csrc", "device")
ORACLE_DIR = os.path.join(ROOT, "oracle")
BUILD = os.path.join(ROOT, "_build")
INCLUDE = os.path.join(ROOT, "include")

LIB = os.path.join(PKG, "libpastix_amd.so")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = os.environ.get("PASTIX_AMD_ARCH", "gfx950")

CFLAGS = ["-O2", "-g", "-fPIC", "-std=gnu11", "-Wall", "-Wno-unused-function",
          "-I" + INCLUDE, "-pthread"]
HIPFLAGS = ["-O3", "-fPIC", "-std=c++17", "--offload-arch=" + ARCH,
            "-munsafe-fp-atomics",
            # packed FP32 instructions are kept out of the device code
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I" + INCLUDE, "-I" + DEV_DIR]


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + proc.stdout)
        raise RuntimeError("build step failed: %s" % os.path.basename(cmd[-1]))
    return proc.stdout


def _obj(src):
    return os.path.join(BUILD, os.path.basename(src) + ".o")


def build_library(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    host_srcs = sorted(os.path.join(HOST_DIR, f) for f in os.listdir(HOST_DIR) if f.endswith(".c"))
    dev_srcs = sorted(os.path.join(DEV_DIR, f) for f in os.listdir(DEV_DIR) if f.endswith(".hip"))
    jobs = []
    for s in host_srcs:
        jobs.append(["gcc"] + CFLAGS + ["-c", "-o", _obj(s), s])
    for s in dev_srcs:
        jobs.append([HIPCC] + HIPFLAGS + ["-c", "-o", _obj(s), s])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for out in ex.map(_run, jobs):
            if verbose and out.strip():
                print(out)
    objs = [_obj(s) for s in host_srcs + dev_srcs]
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs +
         ["-lpthread", "-lm", "-ldl"])
    os.replace(tmp, LIB)
    return LIB


def build_oracle(verbose=False):
    srcs = sorted(os.path.join(ORACLE_DIR, f) for f in os.listdir(ORACLE_DIR) if f.endswith(".c"))
    if not srcs:
        return None
    tmp = ORACLE_LIB + ".tmp"
    _run(["gcc", "-O3", "-march=x86-64-v3", "-fPIC", "-shared", "-std=gnu11", "-pthread",
          "-I" + INCLUDE, "-o", tmp] + srcs + ["-lm", "-ldl"])
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


def build(verbose=False):
    lib = build_library(verbose)
    orc = build_oracle(verbose)
    return lib, orc


if __name__ == "__main__":
    print(build(verbose=True))
