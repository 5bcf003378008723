"""Train the warm-start network of the n_x=20 batched family on the GPU.

Generates the training solutions with the batched solver (plain guess and
continuation, Nint=256), trains the network with the project's kernels and
writes the model (default: _model/warmstart_tanh20_m5.npz beside this file,
where bench.py looks for it) plus a .json with the timing and fit figures.

    python collocation-solver-cuda_amd/train_warmstart.py [--n-train 50000] [--out PATH]
"""
import argparse
import importlib.util
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def _package():
    if "colloc_amd" in sys.modules:
        return sys.modules["colloc_amd"]
    spec = importlib.util.spec_from_file_location("colloc_amd", os.path.join(HERE, "__init__.py"),
                                                  submodule_search_locations=[HERE])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["colloc_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--n-train", type=int, default=50000)
    ap.add_argument("--nint", type=int, default=256)
    ap.add_argument("--epochs", type=int, default=200)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    _package()
    from colloc_amd import learned
    _, info = learned.fit_family(a.n_train, a.nint, a.epochs, path=a.out, log=lambda s: print(s, flush=True))
    print(json.dumps(info), flush=True)


if __name__ == "__main__":
    main()
