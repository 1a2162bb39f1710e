"""Builds libjaxfluids_amd.so (HIP, gfx950) in-tree with hipcc.

The library is the product's native core; it is compiled here, on a machine
without a GPU (hipcc cross-compiles), and shipped next to this file.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libjaxfluids_amd.so")
SOURCES = ["jfa_adjoint_ext.hip", "jfa_fused_charprim.hip", "jfa_fused_fsplit.hip", "jfa_fused_charcons.hip", "jfa_fused_prim.hip", "jfa_fused_cons.hip",
           "jfa_batch_f64.hip", "jfa_batch_f32.hip", "jfa_kernels.hip", "jfa_adjoint.hip", "jfa_fused_f64.hip",
           "jfa_fused_f32.hip"]
ARCH = os.environ.get("JFA_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build jaxfluids_amd")


def flags() -> list:
    return [
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        # packed FP32 instructions are disabled for every HIP object of this project
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", os.path.join(ROOT, "include"),
        "-Wno-unused-result",
    ]


def build(verbose: bool = True) -> str:
    """Compile every source (in parallel) into one shared library; returns its path."""
    from concurrent.futures import ThreadPoolExecutor

    cc = hipcc()
    tmp = LIB + ".tmp"
    jobs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        jobs.append(([cc, *flags(), "-c", path, "-o", obj], obj))

    def run(job):
        cmd, _ = job
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)

    with ThreadPoolExecutor(max_workers=min(len(jobs), os.cpu_count() or 8)) as ex:
        list(ex.map(run, jobs))
    objs = [obj for _, obj in jobs]
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for obj in objs:
        os.remove(obj)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
