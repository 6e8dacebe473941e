"""Build liboavbd.so (the HIP step and its C ABI) for gfx950, in-tree."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "liboavbd.so")
SOURCES = [os.path.join(HERE, "csrc", "oavbd.hip")]
DEPS = [os.path.join(HERE, "csrc", f) for f in os.listdir(os.path.join(HERE, "csrc"))] + [
    os.path.join(ROOT, "include", "oavbd.h"), os.path.join(ROOT, "include", "oavbd_fmath.h")]

# -ffp-contract=off: no fused multiply-adds anywhere, so every float result
# is the correctly rounded value of the expression as written (the step's
# bitwise contract with its CPU checker).  Packed FP32 instructions are
# disabled (-packed-fp32-ops).
HIPCC_FLAGS = ["-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc() -> str:
    for p in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.sep not in p or os.path.exists(p):
            return p
    return "hipcc"


def build(force: bool = True, verbose: bool = False) -> str:
    """Compile liboavbd.so; always recompiles unless force is False and the
    library is newer than every source."""
    if not force and os.path.exists(LIB):
        t = os.path.getmtime(LIB)
        if all(os.path.getmtime(p) <= t for p in DEPS if os.path.exists(p)):
            return LIB
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-o", tmp] + SOURCES
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stderr[-6000:])
    if verbose:
        print(r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
