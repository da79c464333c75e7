"""Builds the fp4 HIP library (gfx950) in-tree: fourphonon_amd/libfp4.so, and
its profiling build fourphonon_amd/libfp4prof.so (the same sources compiled
with FP4_PHASE_PROF, which adds per-phase clock counters to the dense kernel;
loaded by the runtime only when FP4_PROFILE_BUILD=1)."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfp4.so")
LIB_PROF = os.path.join(HERE, "libfp4prof.so")
SOURCES = ["fp4_api.hip", "fp4_prefilter.hip", "fp4_rate64.hip", "fp4_dense.hip", "fp4_dense2.hip", "fp4_iterate.hip", "fp4_3ph.hip",
           "fp4_harmonic.hip", "fp4_sample.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off",
    # packed FP32 instructions stay off: some of their forms have returned
    # wrong results on this hardware without raising an error
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def build_lib(verbose: bool = False, jobs: int = 4, profile: bool = False) -> str:
    objdir = os.path.join(HERE, "_build_prof" if profile else "_build")
    os.makedirs(objdir, exist_ok=True)
    objs, procs = [], []
    extra = ["-DFP4_PHASE_PROF"] if profile else []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        cmd = [HIPCC, *FLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
        while len([p for _, p in procs if p.poll() is None]) >= jobs:
            procs[0][1].wait()
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append((src, out))
        elif verbose and out.strip():
            print(out)
    if failed:
        for src, out in failed:
            sys.stderr.write(f"--- {src}\n{out}\n")
        raise RuntimeError("hipcc failed: " + ", ".join(s for s, _ in failed))
    out = LIB_PROF if profile else LIB
    tmp = out + ".tmp"
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs], check=True)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build_lib(verbose="-v" in sys.argv, profile="--profile" in sys.argv))
