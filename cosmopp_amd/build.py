"""Compile the HIP library cosmopp_amd/libcosmopp_amd.so for gfx950.

Usage: python -m cosmopp_amd.build   (or cosmopp_amd.build.build()).
The library is built in-tree so that it travels with the source tree.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libcosmopp_amd.so")
TEST_BIN = os.path.join(HERE, "cmbl_cpp_tests")

HIP_SOURCES = ["cmbl_kernels.hip", "cmbl_map.hip", "cmbl_loops.hip", "cmbl_sampler.hip"]
CXX_SOURCES = ["cmbl_io.cpp", "cmbl_master.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CMBL_OFFLOAD_ARCH", "gfx950")

# FP64 throughout; no fast-math (1e-10 parity).  Packed FP32 is disabled for
# every HIP translation unit.
HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return out


def build(verbose=False):
    objs = []
    obj_dir = os.path.join(HERE, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    for src in HIP_SOURCES:
        obj = os.path.join(obj_dir, src + ".o")
        _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in CXX_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(obj_dir, src + ".o")
        _run([HIPCC] + HIP_FLAGS + ["-c", path, "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    # C++ tests of the host interface (run by pytest when a GPU is present)
    test_src = os.path.join(CSRC, "cmbl_cpp_tests.cpp")
    if os.path.exists(test_src):
        _run([HIPCC, "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC, test_src,
              "-o", TEST_BIN, "-L" + HERE, "-lcosmopp_amd", "-Wl,-rpath," + HERE])
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
