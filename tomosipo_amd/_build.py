"""Compile ``libtsx.so`` in-tree with hipcc for gfx950."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(CSRC, "libtsx.so")
SOURCES = ["tsx.hip"]
HEADERS = ["tsx_plan.h", os.path.join("..", "..", "include", "tsx.h")]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 instructions stay off in every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    # roctx ranges around the entry points (rocprofv3 --marker-trace)
    "-L/opt/rocm/lib", "-lrocprofiler-sdk-roctx", "-Wl,-rpath,/opt/rocm/lib",
]


def hipcc():
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    return os.path.join(rocm, "bin", "hipcc")


def build(verbose=False):
    """Compile every HIP source into ``csrc/libtsx.so`` (always recompiles)."""
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-o", tmp, *[os.path.join(CSRC, s) for s in SOURCES]]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{' '.join(cmd)}\n{res.stderr}")
    if verbose and res.stderr:
        print(res.stderr)
    os.replace(tmp, LIB)
    return LIB


PEAKS_SRC = os.path.join(CSRC, "peaks.hip")
PEAKS_BIN = os.path.join(CSRC, "peaks")


def build_peaks():
    """Compile the peak-rate microbenchmarks (csrc/peaks.hip) into ``csrc/peaks``."""
    flags = [f for f in HIPCC_FLAGS if f not in ("-shared", "-fPIC") and not f.startswith(("-L", "-l", "-Wl"))]
    cmd = [hipcc(), *flags, "-o", PEAKS_BIN, PEAKS_SRC]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{' '.join(cmd)}\n{res.stderr}")
    return PEAKS_BIN


if __name__ == "__main__":
    import sys

    if len(sys.argv) > 1 and sys.argv[1] == "peaks":
        print(build_peaks())
    else:
        print(build(verbose=True))
