"""Build the HIP library (libcbns3d_amd.so) and the host facade in-tree.

    python -m cbns3d_cart_mpi_amd.build          # or: python cbns3d_cart_mpi_amd/build.py

Compiles for gfx950 only.  -ffp-contract=off keeps every FP64 operation
separately rounded (no FMA contraction), which the plain path needs to round
exactly as the reference does; packed FP32 generation is disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libcbns3d_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-ffp-contract=off", "-fno-fast-math",
    "-fPIC", "-Wall", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

SOURCES = ["cbns_kernels.hip", "cbns_kernels_fma.hip", "cbns_kernels_fast.hip", "cbns_kernels_zfuse.hip",
           "cbns_capi.cpp", "cbns_batch.hip", "cbns_batch_capi.cpp", "cbns_spectra.hip"]
# the contracted FP64 form's face kernels are the only code built with FMA
# contraction (the tolerance-bound form writes its fused multiply-adds out)
CONTRACT = {"cbns_kernels_fma.hip": ["-ffp-contract=fast"]}


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_lib(verbose_resources=False):
    objs = []
    for s in SOURCES:
        src = os.path.join(CSRC, s)
        obj = os.path.join(CSRC, s.rsplit(".", 1)[0] + ".o")
        extra = ["-Rpass-analysis=kernel-resource-usage"] if verbose_resources and s.endswith(".hip") else []
        flags = [f for f in HIP_FLAGS if not (s in CONTRACT and f.startswith("-ffp-contract"))]
        run([HIPCC, *flags, *CONTRACT.get(s, []), *extra, "-x", "hip", "-c", src, "-o", obj])
        objs.append(obj)
    run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", LIB, "-L/opt/rocm/lib", "-lhipfft",
         "-Wl,-rpath,/opt/rocm/lib"])
    return LIB


HOST = os.path.join(HERE, "host")
BIN = os.path.join(HERE, "bin")


def build_host():
    """The C++17 host facade (Block3d / SimulationParams / read_input /
    output_vtk_xml surface) and the cbns3d_amd driver, linked to the library."""
    srcs = sorted(os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".cpp")) if os.path.isdir(HOST) else []
    if not srcs:
        return None
    os.makedirs(BIN, exist_ok=True)
    exe = os.path.join(BIN, "cbns3d_amd")
    run(["g++", "-O2", "-std=c++17", "-Wall", "-pthread", "-I" + os.path.join(ROOT, "include"), "-I" + HOST, *srcs,
         "-o", exe, "-L" + HERE, "-lcbns3d_amd", "-Wl,-rpath,$ORIGIN/.."])
    return exe


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    build_lib(verbose_resources="--resources" in argv)
    build_host()


if __name__ == "__main__":
    main()
