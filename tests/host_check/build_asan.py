"""Recipe of the host-side sanitizer build of libcolloc (used by
tests/test_host_sanitizer.py): the host code of colloc_runtime.cpp with
AddressSanitizer and UndefinedBehaviorSanitizer, the device code
uninstrumented."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
PKG = os.path.join(ROOT, "collocation-solver-cuda_amd")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def build_library_sanitized(out_dir, verbose=False):
    """The host code of colloc_runtime.cpp built with AddressSanitizer and
    UndefinedBehaviorSanitizer (-O1 -g, frame pointers kept) into
    out_dir/libcolloc_asan.so, for the host-side sanitizer check; the device
    code is not instrumented.  A program using it links the sanitizer
    runtimes itself (clang++ -fsanitize=address,undefined)."""
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "libcolloc_asan.so")
    cmd = [HIPCC, "-shared", "-fPIC", "-O1", "-g", "-std=c++17", "--offload-arch=gfx950",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
           "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-fno-sanitize-recover=undefined",
           "-Wno-unused-value", "-Wno-unused-result", "-I", os.path.join(ROOT, "include"),
           os.path.join(PKG, "csrc", "colloc_runtime.cpp"), "-o", out, "-lhiprtc"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("building libcolloc_asan.so failed:\n%s\n%s" % (" ".join(cmd), res.stderr))
    if verbose and res.stderr:
        print(res.stderr)
    return out


if __name__ == "__main__":
    print(build_library_sanitized(sys.argv[1] if len(sys.argv) > 1 else os.path.join(HERE, "_build"), verbose=True))
