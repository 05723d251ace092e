"""Builds libpar2gpu.so in-tree with hipcc for gfx950.

    python par2rust_amd/build.py

Each source compiles to an object under par2rust_amd/_build/, then the
objects link into par2rust_amd/libpar2gpu.so, where the ctypes loader in
par2rust_amd/__init__.py looks for it.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "libpar2gpu.so")

# (source, language)
SOURCES = [("p2g_kernels.hip", "hip"), ("p2g_abi.hip", "hip"), ("p2g_batch.cpp", "hip"),
           ("p2g_serve.cpp", "hip"), ("p2g_tune.cpp", "hip"), ("p2g_plan.cpp", "hip"),
           ("p2g_noalloc.hip", "hip"),
           ("par2_create.cpp", "hip"), ("par2_scheme.cpp", "hip")]

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result", "-Wno-unused-value"]
HIP_FLAGS = [
    "--offload-arch=gfx950",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (res.stdout + res.stderr).splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if res.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build step failed: " + " ".join(cmd))
    if verbose and out:
        sys.stderr.write(out + "\n")


def build(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src, lang in SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        flags = COMMON + (HIP_FLAGS if lang == "hip" else ["-D__HIP_PLATFORM_AMD__"])
        _run([hipcc, *flags, "-x", lang, "-c", os.path.join(CSRC, src), "-o", obj], verbose)
        objs.append(obj)
    tmp = OUT + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp, "-lpthread"],
         verbose)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
