"""Build libcylpic.so (the gfx950 HIP library behind include/cylpic.h).

    python -m cylindrical_epoch_amd.build

Compiles every HIP source for gfx950 with hipcc into
cylindrical_epoch_amd/libcylpic.so.  Compiles every time it is called.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcylpic.so")
SOURCES = ["api.cpp", "fields.hip", "push.hip", "push_mixed.hip", "sort.hip", "diag.hip"]
# the mixed-precision push contracts multiply-adds; everything else keeps the
# reference's operation order
CONTRACT = {"push_mixed.hip": "-ffp-contract=fast"}
ARCH = "gfx950"


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def flags():
    return [
        f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
        # FP64 path: the reference's operation order, no fused multiply-add
        "-ffp-contract=off",
        # no packed FP32 VALU ops anywhere
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
        "-I", os.path.join(HERE, "..", "include"),
    ]


def compile_one(src):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"]
    fl = flags()
    if src in CONTRACT:
        fl = [CONTRACT[src] if f == "-ffp-contract=off" else f for f in fl]
    cmd = [hipcc()] + fl + lang + ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return out


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=4) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
