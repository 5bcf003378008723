"""MI355X-native collocation BVP-DAE solver (drop-in for UW-OCP/Collocation-Solver-CUDA).

The directory name is not a Python identifier, so callers load it under the
name ``colloc_amd`` (see ``load_package`` in bench.py / tests/conftest.py).
"""
__version__ = "0.1"
