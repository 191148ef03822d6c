"""Builds librvf_hip.so (gfx950) in-tree with hipcc."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "librvf_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
    # FP64 parity: no FMA contraction anywhere unless a block opts in
    "-ffp-contract=off",
    # packed FP32 math stays off (wrong results seen on this pool)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
SOURCES = [os.path.join(CSRC, "rvf_kernels.hip"), os.path.join(CSRC, "rvf_host.cpp"), os.path.join(CSRC, "rvf_server.cpp"),
           os.path.join(CSRC, "rvf_meshio.cpp")]
LIBS = ["-lz"]  # compressed VTK data


def build_library(verbose=False):
    cmd = [HIPCC] + FLAGS + ["-x", "hip"] + SOURCES + ["-o", LIB + ".tmp"] + LIBS
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    if verbose and r.stderr:
        print(r.stderr)
    os.replace(LIB + ".tmp", LIB)
    return LIB
