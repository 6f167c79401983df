"""Builds lib21cm_hip.so (the cm21 C ABI) for gfx950 with hipcc, in-tree.

Usage: python 21cmpy_amd/build.py [-v]
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib21cm_hip.so")

SOURCES = [
    "cm21_core.hip",
    "cm21_bubbles.hip",
    "cm21_fused.hip",
    "cm21_batched.hip",
    "cm21_unet.hip",
    "cm21_init.hip",
    "cm21_perturb.hip",
    "cm21_deltat.hip",
]

# -packed-fp32-ops off: the compiler emits no packed FP32 add/mul/fma, whose
#   results have been unreliable on this hardware.
# -ffp-contract=off: the parity-critical float32 sequences must not be fused.
# -munsafe-fp-atomics: hardware float atomics on coarse-grained device memory.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=off", "-munsafe-fp-atomics",
    "-Wall", "-Wno-unused-result",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    return "hipcc"


def build(verbose=False):
    """Compile every source to an object and link the shared library."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        cmd = [hipcc(), *HIPCC_FLAGS, *inc, "-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp,
           "-L/opt/rocm/lib", "-lrocfft", "-Wl,--no-undefined"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
