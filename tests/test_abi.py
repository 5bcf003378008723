"""The C ABI: every entry point include/colloc.h declares is exported by
libcolloc.so and bound by the ctypes layer; no compute call is made (no GPU
is needed)."""
import ctypes
import os
import re

import pytest

from colloc_amd import build, native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "colloc.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cs_[a-z_0-9]+)\s*\(", text)))


@pytest.fixture(scope="module")
def libpath():
    if not os.path.exists(build.LIB_PATH):
        build.build_library()
    return build.LIB_PATH


def test_header_and_bindings_agree():
    names = header_functions()
    assert len(names) >= 20
    assert names == sorted(native.SIGNATURES), set(names) ^ set(native.SIGNATURES)


def test_library_exports_every_symbol(libpath):
    L = ctypes.CDLL(libpath)
    for name in header_functions():
        assert hasattr(L, name), name


def test_host_only_calls(libpath):
    L = native.lib()
    assert L.cs_version().decode().startswith("colloc")
    assert isinstance(L.cs_last_error(), (bytes, type(None)))


def test_null_arguments_are_rejected(libpath):
    L = native.lib()
    out = ctypes.c_void_p()
    assert L.cs_create(None, ctypes.byref(out)) == native.CS_E_ARG
    assert L.cs_set_mesh(None, None, 5) == native.CS_E_ARG
    assert L.cs_newton(None, 0.1, 1e-8, 10, None, None, None, 0) == native.CS_E_ARG


def test_invalid_sizes_are_rejected_before_allocation(libpath):
    """Shapes no kernel set accepts are refused before any device call (ny
    above 64 is not among them: the large-block kernels take it)."""
    L = native.lib()
    out = ctypes.c_void_p()
    for bad in (dict(ny=4, nz=1, np=4, m=1, n_max=100, nc=0), dict(ny=0, nz=1, np=4, m=4, n_max=100, nc=0),
                dict(ny=128, nz=32, np=128, m=4, n_max=1, nc=0), dict(ny=4, nz=-1, np=4, m=4, n_max=100, nc=0)):
        d = native.CSDims(**bad)
        assert L.cs_create(ctypes.byref(d), ctypes.byref(out)) == native.CS_E_SHAPE, bad


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(native, "_lib", None)
    monkeypatch.setattr(native, "LIB_PATH", "/nonexistent/libcolloc.so")
    with pytest.raises(ImportError):
        native.lib()
