"""CPU baseline of the benchmark workloads.  Measurement infrastructure: only
bench.py's CPU leg and oracle/calibrate_reference.py use this module.

The reference runs one (T, H) point per process (win.py:81-103, a
multiprocessing pool over the temperature list); so does this: P worker
processes, one point each, each timing one MCMainFunction call on the
workload's own lattice.  The engine is

  "reference"  the reference's own C engine, mcsolver/isingLib.c or
               mcsolver/heisenbergLib.c compiled from its sources into
               oracle/_ref by `make -C oracle ref` (setup.py:7-9 flags), where
               that build exists;
  "port"       the FP64 restatement of the same algorithm (libmcref, glibc
               rand(), random-sequential Metropolis, every observable), for
               XY (xyLib.c is not built: its site pick overflows int,
               xyLib.c:324,384) and wherever oracle/_ref is absent.

Workers must be forked before the parent initialises the GPU.
"""
from __future__ import annotations

import ctypes
import glob
import importlib.util
import multiprocessing as mp
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# workload -> (model, reference module, temperature range of its points)
WORKLOADS = {
    "xy_ladder_scan": ("xy", None, (0.7, 1.2)),
    "ising_square32": ("ising", "isinglib", (1.5, 3.5)),
    "heis_honeycomb64": ("heis", "heisenberglib", (0.1, 2.0)),
    "heis_triangular_dmi128": ("heis", "heisenberglib", (0.05, 1.0)),
    "heis_cubic64": ("heis", "heisenberglib", (1.2, 1.7)),
}


def _lattice(workload: str):
    from mcsolver_amd import lattice as lat
    from mcsolver_amd import workloads as W
    if workload == "xy_ladder_scan":
        return lat.square(32, J=-1.0), "XY square L=32 (the scan's L=32 ladder)"
    if workload == "ising_square32":
        return lat.square(32, J=-1.0, ising=True), "Ising square L=32"
    if workload == "heis_honeycomb64":
        return W.honeycomb_full_tensor(64), "Heisenberg honeycomb 64x64x2, full-tensor bonds"
    if workload == "heis_triangular_dmi128":
        return W.triangular_neel_dmi(128), "Heisenberg triangular 128x128, Neel DMI"
    if workload == "heis_cubic64":
        return lat.cubic(64, J=-1.0), "Heisenberg cubic 64^3"
    raise ValueError(workload)


def reference_path(name: str):
    """Path of the reference's compiled engine in oracle/_ref, or None."""
    paths = glob.glob(os.path.join(HERE, "_ref", name + ".*.so"))
    return paths[0] if paths else None


def reference_module(name: str):
    """The reference's compiled engine from oracle/_ref (its module
    initialiser prints a line), or None."""
    path = reference_path(name)
    if path is None:
        return None
    spec = importlib.util.spec_from_file_location(name, path)
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def _args(model, Lt, T, h, nsweep):
    from mcsolver_amd import lattice as lat
    if model == "ising":
        return lat.ising_args(Lt, T, h, 0, nsweep)
    return lat.on_args(Lt, T, h, 0, nsweep)


def run_point(model: str, kind: str, args: tuple, seed: int) -> float:
    """Seconds of one MCMainFunction call (the call only, not its inputs)."""
    if kind == "reference":
        mod = reference_module(WORKLOAD_MODULE[model])
        libc = ctypes.CDLL("libc.so.6")
        libc.srand(seed)
        t0 = time.perf_counter()
        mod.MCMainFunction(*args)
        return time.perf_counter() - t0
    from oracle import mcref
    t0 = time.perf_counter()
    if model == "ising":
        mcref.ising_run(args, seed=seed)
    else:
        mcref.on_run(2 if model == "xy" else 3, args, seed=seed)
    return time.perf_counter() - t0


WORKLOAD_MODULE = {"ising": "isinglib", "heis": "heisenberglib", "xy": None}

_LAT = {}


def _quiet():
    """Pool initialiser: the reference's engines printf progress lines; a
    worker's stdout goes to /dev/null so that they never mix with the
    caller's output."""
    dn = os.open(os.devnull, os.O_WRONLY)
    os.dup2(dn, 1)
    os.close(dn)


def _worker(job):
    workload, model, kind, T, h, nsweep, seed = job
    Lt = _LAT[workload]
    dt = run_point(model, kind, _args(model, Lt, T, h, nsweep), seed)
    return Lt.n * nsweep, dt


def default_kind(workload: str) -> str:
    model, mod, _ = WORKLOADS[workload]
    return "reference" if mod is not None and reference_path(mod) is not None else "port"


def baseline(workload: str, target_s: float = 10.0, procs: int | None = None, kind: str | None = None) -> dict:
    """Aggregate spin updates/s of P concurrent points (P = usable cores, at
    most 16) of the workload's lattice, each running about target_s seconds
    (its sweep count calibrated on one short call first)."""
    model, _, (t0, t1) = WORKLOADS[workload]
    kind = kind or default_kind(workload)
    Lt, what = _lattice(workload)
    _LAT[workload] = Lt
    P = procs or min(16, len(os.sched_getaffinity(0)))
    temps = [t0 + (t1 - t0) * k / max(P - 1, 1) for k in range(P)]
    with mp.get_context("fork").Pool(P, initializer=_quiet) as pool:
        # calibration call: sweeps per second at the middle temperature
        ns0 = max(1, int(2e5 // Lt.n))
        upd0, dt = pool.apply(_worker, ((workload, model, kind, 0.5 * (t0 + t1), 0.0, ns0, 1),))
        nsweep = max(1, int(target_s * ns0 / max(dt, 1e-9)))
        jobs = [(workload, model, kind, T, 0.0, nsweep, 100 + k) for k, T in enumerate(temps)]
        w0 = time.perf_counter()
        res = pool.map(_worker, jobs, chunksize=1)
        wall = time.perf_counter() - w0
    upd = sum(r[0] for r in res)
    busy = sum(r[1] for r in res)
    src = {"reference": "the reference's own %s.c compiled from its sources (oracle/_ref)" %
                        {"ising": "isingLib", "heis": "heisenbergLib"}.get(model, "?"),
           "port": "the FP64 restatement of %sLib.c MCMainFunction (oracle, glibc rand)" %
                   {"ising": "ising", "heis": "heisenberg", "xy": "xy"}[model]}[kind]
    return {"value": upd / wall, "unit": "spin updates/s", "cores": P, "kind": kind,
            "sample": "%s, %d temperatures in [%g, %g], %d sweeps each, one process per point, random-sequential "
                      "Metropolis with every observable; %s; %.1f s wall" % (what, P, t0, t1, nsweep, src, wall),
            "per_core": upd / busy}
