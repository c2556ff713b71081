"""Build recipe of the engine: compiles the HIP library for gfx950 in-tree.

    python -m lbffs_amd.build          # product library lbffs_amd/liblbffs.so

The product is one shared library, ``lbffs_amd/liblbffs.so``, holding the C ABI of
``include/lbffs.h`` (device step, HIP kernels), ``include/lbffs_geom.h`` (geometry
kernels), ``include/lbffs_case.h`` (host case setup) and ``include/lbffs_output.h``
(output and restart files).  Host case code is compiled by g++ with FMA contraction off so that its float
results match the reference host; device code is compiled by hipcc for gfx950 only.
"""
from __future__ import annotations

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "lbffs_amd")
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(PKG, "liblbffs.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -packed-fp32-ops: keep the compiler from forming v_pk_*_f32 (some of their forms have
#   given wrong results on this hardware now and then).
# -ffp-contract=on: multiply-adds are fused within one expression only, as the source
#   writes them, so every instantiation of a kernel body (two-lattice and in-place
#   streaming, single and batched lattices) rounds alike; fusing across statements is
#   left to the backend's choices per kernel otherwise.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=on",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out)


# kernels.hip is compiled once per (lattice, storage form, single/batch): the step kernels'
# instantiations build in parallel
KERNEL_UNITS = [(q, st, batch) for q in (19, 27) for st in range(4) for batch in (0, 1)]


def build(verbose: bool = True, jobs: int = 0) -> str:
    from concurrent.futures import ThreadPoolExecutor

    os.makedirs(BUILD, exist_ok=True)
    steps = []
    host_objs = []
    for src in ("lbffs_case.cpp", "lbffs_output.cpp"):
        o = os.path.join(BUILD, src.replace(".cpp", ".o"))
        steps.append(["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall",
                      "-c", os.path.join(CSRC, src), "-o", o])
        host_objs.append(o)
    objs = []
    for q, st, batch in KERNEL_UNITS:
        o = os.path.join(BUILD, f"kernels_q{q}_s{st}_b{batch}.o")
        steps.append([HIPCC, *HIP_FLAGS, f"-DKERN_Q={q}", f"-DKERN_STORE={st}", f"-DKERN_BATCH={batch}",
                      "-c", os.path.join(CSRC, "kernels.hip"), "-o", o])
        objs.append(o)
    for src in ("lbffs_api.hip", "lbffs_probe.hip"):
        o = os.path.join(BUILD, src.replace(".hip", ".o"))
        steps.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    n = jobs or min(16, os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=n) as ex:
        for f in [ex.submit(_run, cmd, verbose) for cmd in steps]:
            f.result()
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, *host_objs, "-o", tmp], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
