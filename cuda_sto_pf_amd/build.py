"""Build libpfcore.so (HIP, gfx950) and the main_pf CLI in-tree.

Plain hipcc invocations, one object per translation unit, compiled in
parallel.  Usage: ``python -m cuda_sto_pf_amd.build [--jobs N]``.
"""
import argparse
import os
import shutil
import subprocess
import sys
import time
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpfcore.so")
CLI = os.path.join(HERE, "bin", "main_pf")

HIP_SOURCES = ["pf_engine_f64.hip", "pf_engine_f32.hip", "pf_runtime.hip", "pf_nn.hip", "pf_batch.hip"]
CXX_SOURCES = ["pf_params.cpp", "pf_checkpoint.cpp"]
CLI_SOURCES = ["main_pf.cpp"]
PASS_GROUPS = 8   # z-pass operator groups 0-4, 6 and 7, the column passes (5), pf_pass_inst.hpp

ARCH = os.environ.get("PF_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]
# packed-FP32 instructions stay disabled: some of their forms have given wrong
# results on this hardware without an error
HIP_FLAGS = ["-x", "hip", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-munsafe-fp-atomics"]


def _run(cmd):
    t = time.time()
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")
    return time.time() - t


def build(jobs=None, verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(os.path.dirname(CLI), exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    steps = []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJ, src + ".o")
        steps.append((obj, [HIPCC, *COMMON, *HIP_FLAGS, "-c", "-o", obj, os.path.join(CSRC, src)]))
    # the pass launchers of each group and precision compile in parallel
    # with the engine (pf_pass_inst.hpp)
    for real in ("float", "double"):
        for g in range(PASS_GROUPS):
            obj = os.path.join(OBJ, f"pf_pass_inst_{real}_{g}.o")
            steps.append((obj, [HIPCC, *COMMON, *HIP_FLAGS, f"-DPF_PASS_GROUP={g}",
                                f"-DPF_PASS_REAL_IS_FLOAT={1 if real == 'float' else 0}", "-c", "-o", obj,
                                os.path.join(CSRC, "pf_pass_inst.hip")]))
    for src in CXX_SOURCES:
        obj = os.path.join(OBJ, src + ".o")
        steps.append((obj, ["g++", *COMMON, "-pthread", "-c", "-o", obj, os.path.join(CSRC, src)]))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        times = list(ex.map(lambda s: _run(s[1]), steps))
    if verbose:
        for (obj, _), t in zip(steps, times):
            print(f"[build] {os.path.basename(obj)} {t:.1f}s")
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-pthread", "-o", tmp, *[s[0] for s in steps]])
    os.replace(tmp, LIB)
    cli_srcs = [os.path.join(CSRC, s) for s in CLI_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    if cli_srcs:
        _run(["g++", "-O2", "-std=c++17", "-o", CLI, *cli_srcs, f"-L{HERE}", "-lpfcore",
              f"-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath," + HERE])
    if verbose:
        print(f"[build] wrote {LIB}")
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    a = ap.parse_args()
    build(a.jobs)
