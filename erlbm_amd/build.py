"""Build recipe for the HIP library (gfx950): ``python -m erlbm_amd.build``.

Compiles every ``csrc/*.hip`` / ``csrc/*.cpp`` source into one shared
library ``erlbm_amd/liberlbm.so`` with hipcc, in-tree.  FP contraction is off
so the FP64 path keeps the reference's rounding, and packed FP32 ops are
disabled.
"""
from __future__ import annotations

import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
LIB = os.path.join(HERE, "liberlbm.so")
ARCH = os.environ.get("ERLBM_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-std=c++17", "-O3", "-fPIC",
    f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))


def build(verbose: bool = False, jobs: int = 8) -> str:
    """Compile all sources (always; no staleness check) and link liberlbm.so."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    cc = hipcc()
    procs, objs = [], []
    for src in sources():
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        objs.append(obj)
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmd = [cc, *HIP_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", *lang, "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        if len([p for _, p in procs if p.poll() is None]) >= jobs:
            procs[0][1].wait()
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        # hipcc's host pass warns about the packed-fp32 feature; harmless
        lines = [ln for ln in text.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
        if p.returncode != 0:
            failed.append(src)
            print("\n".join(lines), file=sys.stderr)
        elif verbose and lines:
            print("\n".join(lines))
    if failed:
        raise RuntimeError(f"hipcc failed for {failed}")
    tmp = LIB + ".tmp"
    cmd = [cc, "-shared", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_drivers(verbose)
    return LIB


DRIVERS = os.path.join(HERE, "drivers")
BIN = os.path.join(HERE, "bin")


def build_drivers(verbose: bool = False) -> list:
    """The C++ drivers (drivers/*.cpp): host programs over the C ABI, linked
    against liberlbm.so (found next to bin/ at run time)."""
    os.makedirs(BIN, exist_ok=True)
    cxx = shutil.which("g++") or "g++"
    out = []
    for src in sorted(glob.glob(os.path.join(DRIVERS, "*.cpp"))):
        exe = os.path.join(BIN, os.path.splitext(os.path.basename(src))[0])
        cmd = [cxx, "-std=c++17", "-O2", "-ffp-contract=off", f"-I{INCLUDE}", f"-I{DRIVERS}", src, "-o", exe + ".tmp",
               f"-L{HERE}", "-lerlbm", "-Wl,-rpath,$ORIGIN/.."]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(exe + ".tmp", exe)
        out.append(exe)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
