"""Build recipe of the engine library: hipcc for gfx950, in-tree output."""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libmpm_amd.so")
SOURCES = [os.path.join(HERE, "csrc", f) for f in ("kernels.hip", "mlp.hip", "engine.cpp", "research.cpp")]
HEADERS = [os.path.join(HERE, "csrc", "engine_internal.h"), os.path.join(HERE, "csrc", "physics.h"),
           os.path.join(HERE, "csrc", "mlp.h"), os.path.join(HERE, "csrc", "json.h"),
           os.path.join(HERE, "csrc", "host_math.h"),
           os.path.join(ROOT, "include", "mpm_engine.h")]

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-munsafe-fp-atomics",
    # Packed FP32 (v_pk_*_f32) stays off: scalar FP32 only.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    cmd = [hipcc()] + FLAGS + ["-I", os.path.join(ROOT, "include")] + SOURCES + ["-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
