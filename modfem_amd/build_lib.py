"""Build recipe for libmodfem_hip.so (gfx950).

Host-only C++ (mesh, setup) is compiled with g++/OpenMP; the HIP translation
units with hipcc for gfx950 only.  The shared library is written in-tree next
to this file so that it travels with the repository snapshot.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmodfem_hip.so")
BUILD = os.path.join(HERE, "_build")

HOST_SRCS = ["mf_mesh.cpp", "mf_setup.cpp", "mf_mkb.cpp"]
HIP_SRCS = ["mf_device.hip", "mf_solver.hip", "mf_capi.hip", "mf_lar.hip", "mf_ilu.hip",
            "mf_gs.hip", "mf_dg.hip", "mf_bcrs.hip", "mf_accel.hip", "mf_setup_dev.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Wno-unused-result",
    # packed FP32 instructions stay off for every HIP translation unit
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
HOST_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-fopenmp"]
# translation units whose kernels repeat the reference's rounding step by step:
# no multiply-add contraction anywhere in them (hipcc contracts by default and
# then ignores "#pragma clang fp contract")
FILE_FLAGS = {"mf_bcrs.hip": ["-ffp-contract=off"], "mf_setup_dev.hip": ["-ffp-contract=off"]}


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return r.stdout


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for s in HOST_SRCS:
        o = os.path.join(BUILD, s + ".o")
        _run(["g++"] + HOST_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    for s in HIP_SRCS:
        o = os.path.join(BUILD, s + ".o")
        _run([HIPCC] + HIP_FLAGS + FILE_FLAGS.get(s, []) + ["-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lgomp"])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_test_drivers(verbose=False):
    """the compiled C drivers of the test suite (tests/c): plain C programs
    that reach the library only through the reference's seam headers"""
    root = os.path.dirname(HERE)
    src = os.path.join(root, "tests", "c", "seam_driver.c")
    out = os.path.join(root, "tests", "c", "seam_driver")
    _run(["gcc", "-O2", "-std=c99", "-Wall", "-I" + os.path.join(root, "include"), src,
          "-L" + HERE, "-lmodfem_hip", "-Wl,-rpath," + "$ORIGIN/../../modfem_amd", "-lm",
          "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    build(verbose=True)
