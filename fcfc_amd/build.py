"""Build the native parts of fcfc_amd in-tree.

  fcfc_amd/lib/libfcfc_gpu.so   gfx950 kernels + C ABI + host C (one library)
  fcfc_amd/bin/FCFC_2PT_BOX_GPU drop-in executables (C host over the C ABI)
  fcfc_amd/bin/FCFC_2PT_GPU

Usage: python -m fcfc_amd.build   (or fcfc_amd.build.build()).
hipcc cross-compiles for gfx950, so this runs on a machine without a GPU.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
BIN = os.path.join(HERE, "bin")
INC = os.path.join(ROOT, "include")
BUILD = os.path.join(ROOT, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CC = os.environ.get("CC", "gcc")

# Packed FP32 instructions are disabled for all device code; the FP64 pair
# arithmetic must not be contracted into FMAs other than the explicit ones.
# The pair loop asks for a 2x unroll, which the compiler declines for the
# variants with re-evaluation queues or cross-lane sums (a note, not an error).
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-Wno-pass-failed",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
C_FLAGS = ["-O2", "-std=c99", "-D_POSIX_C_SOURCE=200809L", "-fPIC",
           "-ffp-contract=off", "-Wall", "-Wextra"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in l)
    if out.strip():
        print(out, flush=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}")


def build(verbose=True):
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    os.makedirs(BUILD, exist_ok=True)
    host_o = os.path.join(BUILD, "fcfc_host.o")
    _run([CC] + C_FLAGS + ["-I" + INC, "-c", os.path.join(CSRC, "fcfc_host.c"), "-o", host_o])
    gpu_o = os.path.join(BUILD, "fcfc_gpu.o")
    _run([HIPCC] + HIP_FLAGS + ["-I" + INC, "-I" + CSRC, "-c",
                                os.path.join(CSRC, "fcfc_gpu.hip"), "-o", gpu_o])
    tmp = os.path.join(LIB, "libfcfc_gpu.so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, gpu_o, host_o])
    os.replace(tmp, os.path.join(LIB, "libfcfc_gpu.so"))
    cli = os.path.join(CSRC, "fcfc_cli.c")
    if os.path.exists(cli):
        for name, geom in (("FCFC_2PT_BOX_GPU", "0"), ("FCFC_2PT_GPU", "1")):
            out = os.path.join(BIN, name)
            _run([CC] + C_FLAGS + ["-fopenmp", "-DFCFC_CLI_GEOM=" + geom, "-I" + INC, "-o", out, cli,
                                   "-L" + LIB, "-lfcfc_gpu", "-Wl,-rpath,$ORIGIN/../lib", "-lm"])
    return os.path.join(LIB, "libfcfc_gpu.so")


def build_oracle():
    """The CPU checker (oracle/): its C restatement, and the upstream
    reference binaries when the upstream sources are present."""
    od = os.path.join(ROOT, "oracle")
    _run(["make", "-C", od, "-j4"])
    if os.path.isdir(os.environ.get("FCFC_REFDIR", "/root/reference")):
        _run(["make", "-C", od, "ref", "-j6"])


if __name__ == "__main__":
    build()
    if "--oracle" in sys.argv:
        build_oracle()
