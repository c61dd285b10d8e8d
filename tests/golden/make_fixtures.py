"""Generate replay fixtures from the reference's own compiled C engines.

Runs only where /root/reference exists: `make -C oracle ref` compiles
mcsolver/isingLib.c and mcsolver/heisenbergLib.c (flags of setup.py:7-9) into
oracle/_ref/, this script imports those two extension modules, reseeds the
process-global glibc rand() with srand(seed) (the reference never calls
srand, fact: its RNG state is libc's), calls MCMainFunction on each case of
cases.py and writes the outputs to replay_fixtures.json.  The fixture holds
data only: case name, SHA-256 of the input tuples, the seed and the
reference's outputs.  xyLib.c is not run: its site pick overflows int on
Linux (xyLib.c:324,384).

    python tests/golden/make_fixtures.py
"""
from __future__ import annotations

import contextlib
import ctypes
import glob
import importlib.util
import io
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from cases import CASES, args_digest, case_args  # noqa: E402


def _load(name):
    paths = glob.glob(os.path.join(ROOT, "oracle", "_ref", name + ".*.so"))
    if not paths:
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
        paths = glob.glob(os.path.join(ROOT, "oracle", "_ref", name + ".*.so"))
    spec = importlib.util.spec_from_file_location(name, paths[0])
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def _jsonable(x):
    if isinstance(x, tuple):
        return [_jsonable(v) for v in x]
    return x


def main():
    libc = ctypes.CDLL("libc.so.6")
    mods = {"ising": _load("isinglib"), "heis": _load("heisenberglib")}
    out = {}
    for name, (model, _mk, _kw, seed) in CASES.items():
        args = case_args(name)
        libc.srand(seed)
        with contextlib.redirect_stdout(io.StringIO()):
            res = mods[model].MCMainFunction(*args)
        out[name] = {"model": model, "seed": seed, "args_sha256": args_digest(args), "result": _jsonable(res)}
        print(name, "ok", flush=True)
    with open(os.path.join(HERE, "replay_fixtures.json"), "w") as f:
        json.dump(out, f, indent=0)


if __name__ == "__main__":
    main()
