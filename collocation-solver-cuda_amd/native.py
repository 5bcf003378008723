"""ctypes binding of libcolloc.so (declarations in include/colloc.h).

There is no fallback: if the library is missing the import of the solver
fails with an error that says how to build it.
"""
import ctypes
import os

import numpy as np

from .build import LIB_PATH

c_int, c_double, c_size_t, c_void_p, c_char_p = ctypes.c_int, ctypes.c_double, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_char_p
P_D = ctypes.POINTER(ctypes.c_double)
P_I = ctypes.POINTER(ctypes.c_int)

CS_OK, CS_E_HIP, CS_E_SHAPE, CS_E_SINGULAR, CS_E_ARG, CS_E_COMPILE, CS_E_NOMODULE, CS_E_UNSUPPORTED = range(8)
KERNELS = {"residual": 0, "jacobian": 1, "condense": 2, "babd": 3, "recover": 4}

# every symbol include/colloc.h declares, with (restype, argtypes)
SIGNATURES = {
    "cs_last_error": (c_char_p, []),
    "cs_version": (c_char_p, []),
    "cs_jit_compile": (c_int, [c_char_p, c_char_p, c_char_p, ctypes.POINTER(c_void_p), ctypes.POINTER(c_size_t),
                               ctypes.POINTER(c_void_p)]),
    "cs_free": (None, [c_void_p]),
    "cs_create": (c_int, [c_void_p, ctypes.POINTER(c_void_p)]),
    "cs_destroy": (None, [c_void_p]),
    "cs_load_module": (c_int, [c_void_p, c_void_p, c_size_t]),
    "cs_set_mesh": (c_int, [c_void_p, P_D, c_int]),
    "cs_set_batch_mesh": (c_int, [c_void_p, c_int, P_I, P_D]),
    "cs_set_constants": (c_int, [c_void_p, P_D, c_int]),
    "cs_set_batch_constants": (c_int, [c_void_p, P_D, c_int, c_int]),
    "cs_batch_info": (c_int, [c_void_p, P_I, P_I]),
    "cs_set_state": (c_int, [c_void_p, P_D, P_D, P_D, P_D]),
    "cs_get_state": (c_int, [c_void_p, P_D, P_D, P_D, P_D]),
    "cs_form_initial_input": (c_int, [c_void_p, P_D, P_D, P_D]),
    "cs_form_initial_input_device": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "cs_restart": (c_int, [c_void_p]),
    "cs_compute_f_q": (c_int, [c_void_p, c_double, P_D, P_D, P_D, P_D, P_D]),
    "cs_compute_f_q_noalloc": (c_int, [c_void_p, c_double, P_D]),
    "cs_construct_jacobian": (c_int, [c_void_p, c_double, P_D, P_D, P_D]),
    "cs_reduce_jacobian": (c_int, [c_void_p, P_D, P_D, P_D]),
    "cs_solve_babd": (c_int, [c_void_p, P_D, P_D]),
    "cs_recover_delta_k": (c_int, [c_void_p, P_D, P_D]),
    "cs_set_babd_substitution": (c_int, [c_void_p, c_int]),
    "cs_set_launch": (c_int, [c_void_p, c_int, c_int]),
    "cs_get_launch": (c_int, [c_void_p, c_int, P_I]),
    "cs_set_direction_precision": (c_int, [c_void_p, c_int]),
    "cs_set_condense_kernel": (c_int, [c_void_p, c_int]),
    "cs_set_newton_graph": (c_int, [c_void_p, c_int]),
    "cs_set_refinement": (c_int, [c_void_p, c_int]),
    "cs_set_babd_precision": (c_int, [c_void_p, c_int]),
    "cs_linear_residual": (c_int, [c_void_p, P_D]),
    "cs_refine_direction": (c_int, [c_void_p, c_int]),
    "cs_get_direction": (c_int, [c_void_p, P_D, P_D, P_D, P_D]),
    "cs_newton": (c_int, [c_void_p, c_double, c_double, c_int, P_I, P_D, P_D, c_int]),
    "cs_newton_progress": (c_int, [c_void_p, c_double, c_double, c_int, c_int, c_double, P_I, P_D, P_D, c_int]),
    "cs_newton_stall": (c_int, [c_void_p, c_double, c_double, c_int, c_int, c_int, c_double, P_I, P_D, P_D, c_int]),
    "cs_segment_residual": (c_int, [c_void_p, c_double, c_double, P_D, P_D]),
    "cs_segment_residual_ex": (c_int, [c_void_p, c_double, c_double, c_int, c_int, c_int, P_D, P_D]),
    "cs_recover_solution": (c_int, [c_void_p, P_D]),
    "cs_node_solution_device": (c_int, [c_void_p, ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p),
                                        ctypes.POINTER(c_void_p)]),
    "cs_profile": (c_int, [c_void_p, c_int]),
    "cs_profile_get": (c_int, [c_void_p, c_int, P_D, P_I]),
    "cs_status": (c_int, [c_void_p, P_I]),
    "cs_condense_stats": (c_int, [c_void_p, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong),
                                  ctypes.POINTER(ctypes.c_longlong)]),
}


class CSDims(ctypes.Structure):
    _fields_ = [("ny", c_int), ("nz", c_int), ("np", c_int), ("m", c_int), ("n_max", c_int), ("nc", c_int),
                ("batch", c_int)]


class ColloError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("colloc error %d: %s" % (code, msg))
        self.code = code


CS_LAUNCH_CONDENSE_GRID, CS_LAUNCH_RESIDUAL_BLOCK, CS_LAUNCH_COMPUTE_UNITS, CS_LAUNCH_JACOBIAN_GEMM = 0, 1, 2, 3

_lib = None


def lib():
    """Load libcolloc.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("%s is missing: build it with __graft_entry__.build() "
                              "(or python collocation-solver-cuda_amd/build.py)" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != CS_OK:
        raise ColloError(rc, lib().cs_last_error().decode(errors="replace"))


def iptr(a):
    """Pointer to a C-contiguous int32 array."""
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(P_I)


def ptr(a):
    """Pointer to a C-contiguous float64 array (None passes NULL)."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(P_D)


def jit_compile(source, name, include_dir):
    code = c_void_p()
    size = c_size_t()
    log = c_void_p()
    rc = lib().cs_jit_compile(source.encode(), name.encode(), include_dir.encode(), ctypes.byref(code),
                              ctypes.byref(size), ctypes.byref(log))
    text = ctypes.string_at(log.value).decode(errors="replace") if log.value else ""
    if log.value:
        lib().cs_free(log)
    check(rc)
    data = ctypes.string_at(code.value, size.value)
    lib().cs_free(code)
    return data, text
