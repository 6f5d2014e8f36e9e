"""Builds the HIP library libg5.so (gfx950) in-tree.

Every device translation unit is compiled with -ffp-contract=off (the FP64 step must keep the
separate roundings of the host restatement) and with packed FP32 instructions disabled.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libg5.so")
H5LIB = os.path.join(HERE, "libg5h5.so")
# HDF5 1.10 C library of the image (headers, libhdf5 and h5dump under /opt/conda)
HDF5_PREFIX = os.environ.get("G5_HDF5_PREFIX", "/opt/conda")

HIP_SOURCES = ["g5_step2d.hip", "g5_step2d_fast.hip", "g5_step2d_mixed.hip", "g5_step3d.hip", "g5_step3d_fast.hip",
               "g5_step3d_mixed.hip", "g5_step3d_amr.hip",
               "g5_step3d_amr_fast.hip", "g5_aux.hip", "g5_refine.hip", "g5_runtime.hip"]
CXX_SOURCES = ["g5_init.cpp", "g5_specs.cpp", "g5_serve.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("G5_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # FP32 division/sqrt of the mixed-precision path use the hardware reciprocal and square root
    # (FP64 division and sqrt stay IEEE correctly rounded)
    "-fno-hip-fp32-correctly-rounded-divide-sqrt",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall"]
# per-file: the 3-D z-march runs one wavefront per SIMD in FP64, where the ILP-oriented scheduler
# hides more latency (measured 2.6% faster on the 256^3 FP64 step, neutral on the mixed one; the
# 2-D step, at several wavefronts per SIMD, is 10-21% slower with it, so it keeps the default)
# The reassociated FP64 instances (G5_PREC_FP64_FAST) also let the compiler contract a*b+c into
# FMA (held to the same relative L1 bar against the bit-exact path as the rest of that path), within
# a source expression only (-ffp-contract=on): contraction across statements (=fast) depends on the
# code around each inlined copy of a device function, so the z solve of the split and the in-place
# z-solve kernels could round differently and a request's bits would depend on the launch geometry
# (seen with the mixed kernels); per expression, every copy rounds alike (and the reassociated 3-D
# step was 1.4% faster this way, the mixed one 3.7% slower than with =fast).
FILE_FLAGS = {"g5_step3d.hip": ["-mllvm", "-amdgpu-sched-strategy=max-ilp"],
              "g5_step3d_fast.hip": ["-mllvm", "-amdgpu-sched-strategy=max-ilp", "-ffp-contract=on"],
              "g5_step3d_amr.hip": ["-mllvm", "-amdgpu-sched-strategy=max-ilp"],
              "g5_step3d_amr_fast.hip": ["-mllvm", "-amdgpu-sched-strategy=max-ilp", "-ffp-contract=on"],
              "g5_step2d_fast.hip": ["-ffp-contract=on"],
              # the mixed-precision instances may contract too, per expression (held to their accuracy
              # bar against the FP64 oracle); the 3-D ones keep the 3-D file's scheduler.  Their FP32 arithmetic flushes
              # denormals (FP64 keeps IEEE denormals): an FP32 quotient is then the hardware reciprocal
              # and a product, and a square root the hardware one, with no range-scaling steps
              # (6-7% faster steps, measured on the 2-D batch and the 3-D mixed case)
              "g5_step2d_mixed.hip": ["-ffp-contract=on", "-fgpu-flush-denormals-to-zero"],
              "g5_step3d_mixed.hip": ["-mllvm", "-amdgpu-sched-strategy=max-ilp", "-ffp-contract=on",
                                      "-fgpu-flush-denormals-to-zero"]}


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    if verbose and r.stdout.strip():
        print(r.stdout)


def build_hdf5(verbose=False):
    """libg5h5.so: HDF5 + XDMF output and restart (host code, kept out of libg5.so)."""
    inc = os.path.join(HDF5_PREFIX, "include")
    libdir = os.path.join(HDF5_PREFIX, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        raise RuntimeError(f"hdf5.h not found under {inc}: set G5_HDF5_PREFIX")
    tmp = H5LIB + ".tmp"
    _run(["g++", *CXX_FLAGS, "-fopenmp", "-shared", "-I", inc, os.path.join(CSRC, "g5_hdf5.cpp"), "-o", tmp,
          "-L", libdir, "-lhdf5", "-Wl,-rpath," + libdir], verbose)
    shutil.move(tmp, H5LIB)
    return H5LIB


def build(verbose=False, extra_hip_flags=()):
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJ, src + ".o")
        jobs.append((obj, [HIPCC, *HIP_FLAGS, *FILE_FLAGS.get(src, []), *extra_hip_flags, "-c", os.path.join(CSRC, src),
                           "-o", obj]))
    for src in CXX_SOURCES:
        obj = os.path.join(OBJ, src + ".o")
        jobs.append((obj, ["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]))
    # translation units in parallel (the 3-D step file dominates)
    with ThreadPoolExecutor(max_workers=min(len(jobs), max(1, min(8, os.cpu_count() or 1)))) as ex:
        list(ex.map(lambda j: _run(j[1], verbose), jobs))
    objs = [o for o, _ in jobs]
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp, *objs], verbose)
    shutil.move(tmp, LIB)
    build_hdf5(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
