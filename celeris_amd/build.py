"""Builds the HIP/C++ engine library in-tree: celeris_amd/libceleris_amd.so (gfx950).

Usage: python -m celeris_amd.build   (or import and call build()).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libceleris_amd.so")
SOURCES = ["kernels.hip", "engine.cpp", "batch.cpp", "terrain.cpp", "capi.cpp"]

# FMA contraction off: every product and sum rounds as the reference's shader text is
# written.  Packed FP32 instructions are disabled for the device.
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def build(verbose=False, out=LIB, defines=()):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = out + ".tmp"
    cmd = [hipcc] + FLAGS + [f"-D{d}" for d in defines] + ["-x", "hip"] + srcs + ["-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    log = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + log)
    if verbose and log.strip():
        print(log)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
