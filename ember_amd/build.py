"""Build the ember_amd HIP library (gfx950) in-tree.

``python -m ember_amd.build`` compiles ``ember_amd/csrc`` into
``ember_amd/libember_amd.so`` with hipcc.  Objects are cached in
``ember_amd/_build`` and rebuilt when a source or header changes.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libember_amd.so")
ARCH = os.environ.get("EMX_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # contract a*b+c into FMA per source expression only, so every kernel that shares the
    # integrator code computes it with the same roundings (the adjoint re-integrates the
    # forward run and relies on reproducing it bit for bit)
    "-ffp-contract=on",
    # packed FP32 instructions stay off: some of their forms have given wrong results on MI355X
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


# Per-file machine-scheduler strategy (LLVM AMDGPU's iterative schedulers), chosen by timing
# the kernels of each file on MI355X with the default and the iterative strategies: the
# QSS stage, the diffusion chunks and transport run faster with iterative-ilp, the BDF and
# the adjoint with iterative-maxocc.  Instruction order only: the arithmetic is unchanged.
FILE_FLAGS = {
    "qss_kernels.hip": ["-mllvm", "-amdgpu-sched-strategy=iterative-ilp"],
    "diffusion_kernels.hip": ["-mllvm", "-amdgpu-sched-strategy=iterative-ilp"],
    "transport_kernels.hip": ["-mllvm", "-amdgpu-sched-strategy=iterative-ilp"],
    "bdf_kernels.hip": ["-mllvm", "-amdgpu-sched-strategy=iterative-maxocc"],
    "adjoint_kernels.hip": ["-mllvm", "-amdgpu-sched-strategy=iterative-maxocc"],
}


def _sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))


def _headers():
    return sorted(glob.glob(os.path.join(CSRC, "*.h")) + glob.glob(os.path.join(ROOT, "include", "*.h")))


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    return out


def build(force: bool = False, jobs: int = 8) -> str:
    os.makedirs(OBJ, exist_ok=True)
    hdr_mtime = max(os.path.getmtime(h) for h in _headers())
    objs, todo = [], []
    for src in _sources():
        obj = os.path.join(OBJ, os.path.basename(src) + ".o")
        objs.append(obj)
        if force or not os.path.exists(obj) or os.path.getmtime(obj) < max(os.path.getmtime(src), hdr_mtime):
            todo.append((src, obj))
    procs = []
    for src, obj in todo:
        cmd = [HIPCC, *FLAGS, *FILE_FLAGS.get(os.path.basename(src), []), "-c", src, "-o", obj + ".tmp"]
        if src.endswith(".cpp"):
            cmd = [HIPCC, *FLAGS, "-x", "hip", "-c", src, "-o", obj + ".tmp"]
        procs.append((subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True), obj, cmd))
        if len(procs) >= jobs:
            _wait(procs.pop(0))
    for p in procs:
        _wait(p)
    if todo or force or not os.path.exists(LIB):
        _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", LIB + ".tmp", *objs])
        os.replace(LIB + ".tmp", LIB)
    return LIB


def _wait(item):
    p, obj, cmd = item
    out, _ = p.communicate()
    if p.returncode != 0:
        out = "\n".join(l for l in out.splitlines() if "packed-fp32-ops" not in l)
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    os.replace(obj + ".tmp", obj)


def build_test_programs() -> str:
    """C++ test driver for the split-step sequence (links the library and the oracle)."""
    src = os.path.join(ROOT, "tests", "cpp", "split_flame.cpp")
    out_dir = os.path.join(ROOT, "tests", "cpp", "_build")
    os.makedirs(out_dir, exist_ok=True)
    exe = os.path.join(out_dir, "split_flame")
    oracle_lib = os.path.join(ROOT, "oracle", "_build")
    _run([HIPCC, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), src, "-o", exe + ".tmp",
          "-L", PKG, "-lember_amd", "-L", oracle_lib, "-loracle", "-fopenmp",
          f"-Wl,-rpath,{PKG}", f"-Wl,-rpath,{oracle_lib}", "-Wl,-rpath,$ORIGIN/../../../ember_amd",
          "-Wl,-rpath,$ORIGIN/../../../oracle/_build"])
    os.replace(exe + ".tmp", exe)
    return exe


if __name__ == "__main__":
    print(build(force="--force" in sys.argv))
