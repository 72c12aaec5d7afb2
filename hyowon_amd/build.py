"""Build libhyowon_pm.so (hand-written HIP for gfx950) in-tree.

    python -m hyowon_amd.build          # or hyowon_amd.build.build()

Compiles every source under hyowon_amd/csrc with hipcc and links
hyowon_amd/libhyowon_pm.so.  Flags:
  -ffp-contract=off     the binary64 paths restate the reference arithmetic
                        operation by operation (no implicit FMA)
  -packed-fp32-ops off  packed FP32 instructions are not emitted
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
LIB = os.path.join(HERE, "libhyowon_pm.so")
ARCH = os.environ.get("HY_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++20", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
          f"-I{CSRC}", f"-I{INCLUDE}", "-Wall", "-Wno-unused-function"]
HIP_ONLY = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
            "-packed-fp32-ops", "-munsafe-fp-atomics"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise FileNotFoundError("hipcc not found")


def build(verbose: bool = False) -> str:
    cc = hipcc()
    with tempfile.TemporaryDirectory(prefix="hy_build_") as tmp:
        objs = []
        jobs = []
        for src in sorted(os.listdir(CSRC)):
            path = os.path.join(CSRC, src)
            obj = os.path.join(tmp, src + ".o")
            if src.endswith(".hip"):
                cmd = [cc, *COMMON, *HIP_ONLY, "-c", path, "-o", obj]
            elif src.endswith(".cpp"):
                cmd = [cc, *COMMON, "-x", "c++", "-c", path, "-o", obj]
            else:
                continue
            objs.append(obj)
            jobs.append(cmd)
        procs = []
        for cmd in jobs:
            if verbose:
                print(" ".join(cmd), flush=True)
            procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        failed = []
        for cmd, p in procs:
            out, _ = p.communicate()
            if p.returncode != 0:
                failed.append((cmd, out.decode(errors="replace")))
            elif verbose and out:
                print(out.decode(errors="replace"))
        if failed:
            msg = "\n".join(f"$ {' '.join(c)}\n{o}" for c, o in failed)
            raise RuntimeError(f"hipcc failed:\n{msg}")
        tmp_lib = LIB + ".tmp"
        link = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp_lib, *objs,
                "-lquadmath", "-ldl"]
        if verbose:
            print(" ".join(link), flush=True)
        subprocess.run(link, check=True)
        os.replace(tmp_lib, LIB)
    build_cpp_dropin(verbose)
    return LIB


def build_cpp_dropin(verbose: bool = False) -> str:
    """Compile the C++ drop-in example/test (tests/cpp) against the library."""
    root = os.path.dirname(HERE)
    src = os.path.join(root, "tests", "cpp", "dropin_force.cpp")
    out = os.path.join(root, "tests", "cpp", "dropin_force")
    if not os.path.exists(src):
        return ""
    cmd = [shutil.which("g++") or "g++", "-O2", "-std=c++20", f"-I{INCLUDE}", src, "-o", out,
           f"-L{HERE}", "-lhyowon_pm", f"-Wl,-rpath,$ORIGIN/../../hyowon_amd"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
