"""Build step for the native library (host runtime of the solver).

libcolloc.so holds no kernels of its own: every kernel is specialised per
problem and compiled with hiprtc at load time (see jit.py).  The library is
built in-tree with hipcc so that it travels with the source tree.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_DIR = os.path.join(HERE, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libcolloc.so")
KERNEL_DIR = os.path.join(HERE, "csrc", "kernels")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def build_library(verbose=False):
    """Compile csrc/colloc_runtime.cpp into lib/libcolloc.so (always rebuilds)."""
    os.makedirs(LIB_DIR, exist_ok=True)
    src = os.path.join(HERE, "csrc", "colloc_runtime.cpp")
    tmp = LIB_PATH + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "-O2", "-std=c++17", "--offload-arch=gfx950",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wno-unused-value", "-Wno-unused-result", "-I", os.path.join(ROOT, "include"), src, "-o", tmp, "-lhiprtc"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("building libcolloc.so failed:\n%s\n%s" % (" ".join(cmd), res.stderr))
    if verbose and res.stderr:
        print(res.stderr)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build_library(verbose=True))
