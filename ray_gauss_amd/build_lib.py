"""Builds the gfx950 shared library ``_build/libraygauss_hip.so`` from ``csrc/`` with hipcc.

Flags:
  --offload-arch=gfx950                         MI355X (CDNA4) only
  -fhip-fp32-correctly-rounded-divide-sqrt      IEEE division / sqrt: the ray grid and the hit
                                                predicate are exact FP32 and must be reproducible
  -fno-gpu-flush-denormals-to-zero              IEEE denormals, same reason
  -munsafe-fp-atomics                           native global FP32 atomic add (device memory is
                                                coarse-grained hipMalloc memory)
  -Xclang -target-feature -Xclang -packed-fp32-ops   no packed-FP32 instructions
"""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libraygauss_hip.so")
SOURCES = ["preprocess.hip", "lbvh.hip", "sah.hip", "trace.hip", "optim.hip", "loss.hip", "densify.hip", "api.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed: {' '.join(cmd)}")
    return r


def build(verbose: bool = False, jobs: int = 4) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    objs = []
    procs = []
    for src in SOURCES:
        obj = os.path.join(OUT_DIR, src.replace(".", "_") + ".o")
        cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd = [HIPCC, *FLAGS, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        objs.append(obj)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
        if verbose:
            print(" ".join(cmd), flush=True)
        while sum(p.poll() is None for _, p in procs) >= jobs:
            procs[0][1].wait()
    failed = False
    for cmd, p in procs:
        out, err = p.communicate()
        if p.returncode != 0:
            failed = True
            sys.stderr.write(out + err)
        elif verbose and err.strip():
            sys.stderr.write(err)
    if failed:
        raise RuntimeError("hipcc compilation failed")
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
