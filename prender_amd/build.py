"""Builds libprt.so (HIP kernels + host C ABI) for gfx950 in-tree.

    python -m prender_amd.build

hipcc cross-compiles without a GPU.  Every HIP source is compiled with
-ffp-contract=off (the geometry kernels reproduce the reference's double
precision operation order) and with packed FP32 instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "_lib")
LIB_PATH = os.path.join(LIB_DIR, "libprt.so")
SOURCES = ["prt_kernels.hip", "prt_stream.hip", "prt_persist.hip", "prt_mlt.hip", "prt_api.cpp", "prt_batch.cpp", "prt_mlt.cpp", "bvh.cpp",
           "scene_file.cpp", "image_io.cpp"]
HEADERS = ["prt_device.h", "prt_kernels.h", "prt_math.h", "prt_trace.h", "prt_shade.h", "prt_stream.h",
           "prt_internal.h", "prt_mlt.h", "bvh.h"]
ARCH = "gfx950"


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: ROCm is required to build libprt")


def build(verbose: bool = False, defines=(), out: str = LIB_PATH) -> str:
    """Compile libprt.so; ``defines`` are extra -D options (tuning variants)."""
    os.makedirs(os.path.dirname(out), exist_ok=True)
    tmp = out + ".tmp"
    cmd = [hipcc(), "-O3", f"--offload-arch={ARCH}", "-std=c++17", "-fPIC", "-shared",
           "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
           "-I", CSRC] + [f"-D{d}" for d in defines] + [os.path.join(CSRC, s) for s in SOURCES] + \
        ["-o", tmp, "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    log = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError("libprt build failed:\n" + log)
    if verbose and log:
        print(log)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
