"""Builds libkxpbd.so for gfx950 in-tree with hipcc (no JIT cache)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libkxpbd.so")
# sources with no device code: compiled for the host only (they use x86
# builtins such as the run-time AVX2 check)
HOST_ONLY = ("kx_serve.cpp",)
SOURCES = ["kx_api.cpp", "kx_builder.cpp", "kx_sched.cpp", "kx_serve.cpp", "kx_plain.hip", "kx_persist.hip", "kx_policy.hip",
           "kx_noalloc.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # no packed-fp32 instructions anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # fused multiply-adds only where one source expression asks for them:
    # device functions inlined into different kernels then round alike
    "-ffp-contract=on",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def _compile(s, verbose):
    src = os.path.join(CSRC, s)
    obj = os.path.join(CSRC, "_build", s + ".o")
    os.makedirs(os.path.dirname(obj), exist_ok=True)
    lang = ["-x", "hip"] if s.endswith(".hip") else []
    if s in HOST_ONLY:
        lang = ["--offload-host-only"]
    cmd = [HIPCC] + FLAGS + lang + ["-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return obj


def build(verbose=False):
    from concurrent.futures import ThreadPoolExecutor
    jobs = max(1, min(len(SOURCES), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
