"""Builds the HIP/C++ engine library in-tree: celeris_amd/libceleris_amd.so (gfx950).

Usage: python -m celeris_amd.build   (or import and call build()).

The kernels and the engine are compiled twice, once per working precision (FP32 in
namespace cel, FP64 in namespace cel64, selected by CEL_FP64); translation units compile
in parallel into celeris_amd/_build/ and are linked into one shared library.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libceleris_amd.so")
BUILD_DIR = os.path.join(HERE, "_build")
# (source, object, extra defines[, extra flags])
UNITS = [
    ("kernels.hip", "kernels.o", []),
    ("kernels.hip", "kernels64.o", ["CEL_FP64"]),
    ("engine.cpp", "engine.o", []),
    ("engine.cpp", "engine64.o", ["CEL_FP64"]),
    ("batch.cpp", "batch.o", []),
    ("terrain.cpp", "terrain.o", []),
    ("capi.cpp", "capi.o", []),
    ("probe.hip", "probe.o", []),
]
SOURCES = sorted({u[0] for u in UNITS})

# FMA contraction off: every product and sum rounds as the reference's shader text is
# written.  Packed FP32 instructions are disabled for the device.
COMPILE_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
FLAGS = COMPILE_FLAGS          # kept for tools that compile single units


def _hipcc():
    return os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def _filter(text):
    return "\n".join(l for l in text.splitlines() if "packed-fp32-ops" not in l)


def build(verbose=False, out=LIB, defines=(), jobs=None):
    os.makedirs(BUILD_DIR, exist_ok=True)
    tag = "" if not defines else "_" + "_".join(d.split("=")[0].lower() for d in defines)

    def compile_unit(unit):
        src, obj, extra = unit[:3]
        flags = unit[3] if len(unit) > 3 else []
        obj_path = os.path.join(BUILD_DIR, obj.replace(".o", tag + ".o"))
        cmd = [_hipcc()] + COMPILE_FLAGS + flags + [f"-D{d}" for d in list(defines) + extra]
        cmd += ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj_path]
        r = subprocess.run(cmd, capture_output=True, text=True)
        return obj_path, r.returncode, _filter(r.stdout + r.stderr)

    with ThreadPoolExecutor(max_workers=jobs or min(len(UNITS), os.cpu_count() or 1)) as ex:
        results = list(ex.map(compile_unit, UNITS))
    logs = []
    for obj_path, rc, log in results:
        if rc != 0:
            raise RuntimeError(f"hipcc failed for {os.path.basename(obj_path)}:\n{log}")
        if log.strip():
            logs.append(log)
    tmp = out + ".tmp"
    cmd = [_hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC"] + [r[0] for r in results] + ["-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + _filter(r.stdout + r.stderr))
    if verbose and logs:
        print("\n".join(logs))
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
