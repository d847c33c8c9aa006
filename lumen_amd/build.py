"""Build liblumen.so (HIP, gfx950) in-tree.

    python -m lumen_amd.build

Every HIP source is compiled with:
  -ffp-contract=off                      the reference's Float32 expression trees
                                         are evaluated without FMA contraction
  -fhip-fp32-correctly-rounded-divide-sqrt  '/' and sqrt correctly rounded
  -fno-gpu-flush-denormals-to-zero       IEEE denormals, as on the host
  -target-feature -packed-fp32-ops       no packed FP32 instructions
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "lumen_amd", "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(ROOT, "lumen_amd", "liblumen.so")

SOURCES = ["liblumen.hip", "capi_legacy.hip", "legacy.hip", "backward.hip", "loss.hip",
           "optim.hip"]
ARCH = os.environ.get("LUMEN_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-Wno-unused-command-line-argument",
    "-I", INCLUDE,
    "-I", CSRC,
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    """Compile every HIP source and link lumen_amd/liblumen.so."""
    objs = []
    os.makedirs(os.path.join(CSRC, "_build"), exist_ok=True)
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(CSRC, "_build", src.replace(".hip", ".o"))
        cmd = [hipcc(), *HIPCC_FLAGS, "-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError(f"hipcc failed on {src}")
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("link of liblumen.so failed")
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> None:
    """Compile the CPU oracle (test infrastructure) and, when the reference
    sources are present, the reference's own C library into oracle/_ref."""
    odir = os.path.join(ROOT, "oracle")
    res = subprocess.run(["make", "-s", "-C", odir, "-B", "all"], capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("oracle build failed")
    if os.path.exists("/root/reference/c_api/src/lumen_c.c"):
        res = subprocess.run(["make", "-s", "-C", odir, "-B", "ref"], capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError("reference C library build failed")


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    build_oracle()
