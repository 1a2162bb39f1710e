"""Builds oracle/c/libjxf_cpu.so, the C/OpenMP restatement (test infrastructure only).

-march=x86-64-v3 (AVX2/FMA) keeps the library runnable on any current x86 host,
including the GPU machines' CPUs, without tuning to this build host.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libjxf_cpu.so")


def build(verbose=True):
    cmd = ["gcc", "-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared", "-std=c11",
           os.path.join(HERE, "jxf_cpu.c"), "-o", LIB + ".tmp", "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build()
