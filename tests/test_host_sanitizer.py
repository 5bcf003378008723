"""Host-side sanitizer build of the C-ABI library (the debug build of the
race-detection / sanitizer row: AddressSanitizer and UndefinedBehaviorSanitizer
on the host code, device code uninstrumented).

tests/host_check/build_asan.py compiles libcolloc's host code with both
sanitizers; tests/host_check/asan_driver.cpp links it with the sanitizer
runtimes and runs the calls that need no GPU: version and error text,
argument and shape checks, two hiprtc compilations of a generated problem
unit (n_x = 4 tanh family member, m = 5), a broken source (compile error and
log), and handle creation, which without a device fails in its first
allocation and releases what it made.  A sanitizer report (heap overflow,
use after free, undefined behaviour) ends the driver with a non-zero status.
Leak detection is off: the HIP runtime keeps allocations until exit.
"""
import os
import subprocess
import sys

import pytest

from colloc_amd import build, emitter, ocp as ocpmod, problems, tables

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "host_check"))
from build_asan import build_library_sanitized   # noqa: E402
ROOT = os.path.dirname(HERE)
CLANG = "/opt/rocm/llvm/bin/clang++"


@pytest.mark.skipif(not os.path.exists(CLANG), reason="no ROCm clang")
def test_host_code_under_address_and_ub_sanitizers(tmp_path):
    lib = build_library_sanitized(str(tmp_path))
    drv = str(tmp_path / "asan_driver")
    cmd = [CLANG, "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
           "-I", os.path.join(ROOT, "include"), os.path.join(HERE, "host_check", "asan_driver.cpp"),
           "-L", str(tmp_path), "-lcolloc_asan", "-Wl,-rpath," + str(tmp_path), "-o", drv]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    inst = problems.batch_family(1, nx=4, nu=2, nint_range=(16, 64))[0]
    unit = emitter.build_unit(emitter.BvpDaeSymbolic(ocpmod.parse_ocp(inst.ocp_text())), 5,
                              tables.collocation_tables(5))
    src = tmp_path / "unit.hip"
    src.write_text(unit)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:exitcode=23",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:exitcode=24")
    import torch   # noqa: F401  (only to learn whether a device is present)
    device = "1" if torch.cuda.device_count() > 0 else "0"
    run = subprocess.run([drv, str(src), build.KERNEL_DIR, device], capture_output=True, text=True, env=env,
                         timeout=600)
    assert run.returncode == 0, run.stdout + run.stderr
    assert run.stdout.strip().endswith("ok"), run.stdout
    assert "ERROR: AddressSanitizer" not in run.stderr and "runtime error" not in run.stderr, run.stderr
