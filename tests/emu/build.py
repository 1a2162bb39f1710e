"""Builds tests/emu/_build/libjfa_emu.so: the library's batch-stage and adjoint sources compiled for the
host with g++ against the emulation header in tests/emu/hip/ (test infrastructure only; the
product library is built by jaxfluids_amd/build.py with hipcc for gfx950)."""
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "jaxfluids_amd", "csrc")
OUT = os.path.join(HERE, "_build", "libjfa_emu.so")
SOURCES = [os.path.join(CSRC, f) for f in ("jfa_batch_f64.hip", "jfa_batch_f32.hip", "jfa_adjoint.hip")]
DEPS = SOURCES + [os.path.join(CSRC, f) for f in ("jfa_batch.h", "jfa_common.h", "jfa_device.h", "jfa_adjoint.h", "jfa_dissipative.h")] + [
    os.path.join(HERE, "emu_runtime.cpp"), os.path.join(HERE, "hip", "hip_runtime.h"), __file__]


def stale() -> bool:
    if not os.path.exists(OUT):
        return True
    t = os.path.getmtime(OUT)
    return any(os.path.getmtime(d) > t for d in DEPS)


def build(force: bool = False) -> str:
    if not force and not stale():
        return OUT
    cxx = shutil.which("g++")
    if cxx is None:
        raise RuntimeError("g++ not found")
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    flags = ["-std=c++17", "-O2", "-g", "-fPIC", "-pthread", "-Wno-unknown-pragmas", "-Wno-unused-result",
             f"-I{os.path.join(HERE)}", f"-I{CSRC}", f"-I{os.path.join(ROOT, 'include')}"]
    objs = []
    for src in SOURCES + [os.path.join(HERE, "emu_runtime.cpp")]:
        obj = os.path.join(os.path.dirname(OUT), os.path.basename(src) + ".o")
        lang = ["-x", "c++"] if src.endswith(".hip") else []
        subprocess.run([cxx, *flags, *lang, "-c", src, "-o", obj], check=True)
        objs.append(obj)
    subprocess.run([cxx, "-shared", "-pthread", "-o", OUT, *objs], check=True)
    return OUT


if __name__ == "__main__":
    print(build(force=True))
