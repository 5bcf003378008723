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
MLP_LIB_PATH = os.path.join(LIB_DIR, "libcolloc_mlp.so")
KERNEL_DIR = os.path.join(HERE, "csrc", "kernels")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def _hipcc_shared(src, out, extra=(), verbose=False):
    os.makedirs(LIB_DIR, exist_ok=True)
    tmp = out + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "-O3", "-std=c++17", "--offload-arch=gfx950",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wno-unused-value", "-Wno-unused-result", "-I", os.path.join(ROOT, "include"), src, "-o", tmp]
    cmd += list(extra)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("building %s failed:\n%s\n%s" % (os.path.basename(out), " ".join(cmd), res.stderr))
    if verbose and res.stderr:
        print(res.stderr)
    os.replace(tmp, out)
    return out


def build_library(verbose=False):
    """Compile csrc/colloc_runtime.cpp into lib/libcolloc.so (always rebuilds)."""
    return _hipcc_shared(os.path.join(HERE, "csrc", "colloc_runtime.cpp"), LIB_PATH, ["-lhiprtc"], verbose)


def build_mlp_library(verbose=False):
    """Compile the learned-warm-start kernels csrc/mlp_kernels.hip into
    lib/libcolloc_mlp.so (always rebuilds)."""
    return _hipcc_shared(os.path.join(HERE, "csrc", "mlp_kernels.hip"), MLP_LIB_PATH, (), verbose)


if __name__ == "__main__":
    print(build_library(verbose=True))
    print(build_mlp_library(verbose=True))
