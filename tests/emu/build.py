"""Builds tests/emu/_build/libjfa_emu.so: the library's sources compiled for the host with g++
against the emulation header in tests/emu/hip/ (test infrastructure only; the product library is
built by jaxfluids_amd/build.py with hipcc for gfx950).  The result exports the library's C ABI."""
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "jaxfluids_amd", "csrc")
OUT = os.path.join(HERE, "_build", "libjfa_emu.so")
SOURCES = [os.path.join(CSRC, f) for f in ("jfa_batch_f64.hip", "jfa_batch_f32.hip", "jfa_adjoint.hip", "jfa_adjoint_ext.hip", "jfa_kernels.hip",
                                            "jfa_fused_f64.hip", "jfa_fused_f32.hip", "jfa_fused_prim.hip",
                                            "jfa_fused_cons.hip", "jfa_fused_charprim.hip", "jfa_fused_charcons.hip",
                                            "jfa_fused_fsplit.hip")]
DEPS = SOURCES + [os.path.join(CSRC, f) for f in ("jfa_batch.h", "jfa_common.h", "jfa_device.h", "jfa_adjoint.h",
                                                  "jfa_dissipative.h", "jfa_fused.h")] + [
    os.path.join(HERE, "emu_runtime.cpp"), os.path.join(HERE, "hip", "hip_runtime.h"), __file__]


def stale() -> bool:
    if not os.path.exists(OUT):
        return True
    t = os.path.getmtime(OUT)
    return any(os.path.getmtime(d) > t for d in DEPS)


def build(force: bool = False) -> str:
    """Builds when stale.  Concurrent callers (pytest-xdist workers) serialise on a lock file and
    the library appears by an atomic rename, so no caller loads a half-written file."""
    import fcntl
    import tempfile

    if not force and not stale():
        return OUT
    cxx = shutil.which("g++")
    if cxx is None:
        raise RuntimeError("g++ not found")
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    with open(OUT + ".lock", "w") as lock:
        fcntl.flock(lock, fcntl.LOCK_EX)
        if not force and not stale():  # built by another process meanwhile
            return OUT
        flags = ["-std=c++17", "-O2", "-g", "-fPIC", "-pthread", "-Wno-unknown-pragmas", "-Wno-unused-result",
                 f"-I{os.path.join(HERE)}", f"-I{CSRC}", f"-I{os.path.join(ROOT, 'include')}"]
        with tempfile.TemporaryDirectory(dir=os.path.dirname(OUT)) as tmp:
            from concurrent.futures import ThreadPoolExecutor

            def compile_one(src):
                obj = os.path.join(tmp, os.path.basename(src) + ".o")
                lang = ["-x", "c++"] if src.endswith(".hip") else []
                subprocess.run([cxx, *flags, *lang, "-c", src, "-o", obj], check=True)
                return obj

            srcs = SOURCES + [os.path.join(HERE, "emu_runtime.cpp")]
            with ThreadPoolExecutor(max_workers=min(len(srcs), 6)) as ex:
                objs = list(ex.map(compile_one, srcs))
            lib = os.path.join(tmp, "lib.so")
            subprocess.run([cxx, "-shared", "-pthread", "-o", lib, *objs], check=True)
            os.replace(lib, OUT)
    return OUT


if __name__ == "__main__":
    print(build(force=True))
