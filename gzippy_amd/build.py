"""In-tree builds of the native libraries.

* ``libgzgpu.so``  — the HIP library (gfx950 kernels + C ABI in ``include/gzgpu.h``)
* ``libgzsynth.so`` — the host-side synthetic content generators

Both land next to this file so they travel with the source tree.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")

HIP_SOURCES = ["gzgpu.hip"]
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 ops stay off (no float math here, but the rule is global)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def gpu_lib_path():
    return os.path.join(HERE, "libgzgpu.so")


def synth_path():
    return os.path.join(HERE, "libgzsynth.so")


def hipcc():
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.exists(c) or c == "hipcc":
            return c
    return "hipcc"


def build_gpu(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    out = gpu_lib_path()
    tmp = out + ".tmp"
    cmd = [hipcc()] + HIP_FLAGS + ["-I", INCLUDE, "-I", CSRC, "-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    return out


def build_synth(verbose=False):
    out = synth_path()
    tmp = out + ".tmp"
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-std=c11", "-o", tmp, os.path.join(CSRC, "synth.c"), "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    return out


def build_all(verbose=False):
    build_synth(verbose)
    build_gpu(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
