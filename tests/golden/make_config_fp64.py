"""FP64 oracle estimates at the benchmark configurations' own lattices: the
reference side of tests/test_gpu_configs.py.

Every run is the oracle restatement of the reference's MCMainFunction
(oracle/ref_glibc.c: random-sequential Metropolis with glibc rand(), FP64,
the reference's observable set) on the positional arguments mcMain.py would
build for that lattice and (T, h); R independent seeds per point.  The
oracle is our own C code (pinned bit for bit to the reference's compiled C
by tests/test_oracle_replay.py), so this runs anywhere the oracle builds.

Sections (per run, the values kept):
  ising_square32   Ising 32x32, J = 1 ferro, h = 0, all-up start, the full
                   schedule of the configuration (2e4 thermal + 5e4 measured
                   sweeps); 4 temperatures of linspace(1.5, 3.5, 16).
                   Keeps the 10 returned floats (spin_i, spin_j, spin_ij,
                   autoCorr, E, E2, E_r, E2_r, U4, spin_tot).
  xy_square_large  XY square L = 64 (3 rungs) and L = 128 (2 rungs) of the
                   ladder linspace(0.7, 1.2, 64).
  xy_square_small  XY square L = 16 and 32, 6 rungs of the same ladder each.
  tri_dmi128       Heisenberg triangular 128x128, Neel DMI 0.3, Dz = -0.05:
                   6 (T, Bz) points of the 32 x 32 grid.
  cubic16          Heisenberg cubic 16^3, J = -1: 3 temperatures of
                   linspace(1.2, 1.7, 128) around T_c = 1.443.
  O(n) sections keep E = tuple[8], E2 = tuple[9] (beta units per site),
  |M| = tuple[22], Q = tuple[26], and <M^2>, <M^4> of the total spin from
  the groups output (one group holding the whole lattice).

Integer draws: --rng=xoshiro (default; writes config_fp64.json) feeds the
restatement from xoshiro256**, --rng=glibc (writes config_fp64_glibc.json)
from glibc rand() as the reference does.  The reference's generator biases
its own averages (glibc rand() is a lagged-Fibonacci generator with lags 3
and 31, and one update consumes 3 to 6 consecutive values): at the
triangular-DMI points the two energies differ by ~9 standard errors, and
tests/test_reference_rng.py keeps that record.  The parity tests compare
against the reference algorithm driven by the sound generator.

    python tests/golden/make_config_fp64.py [--rng=xoshiro|glibc] [section ...]
"""
import json
import multiprocessing as mp
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def out_path(rng):
    return os.path.join(HERE, "config_fp64.json" if rng == "xoshiro" else "config_fp64_%s.json" % rng)

ISING_T_IDX = (5, 6, 7, 10)                    # of linspace(1.5, 3.5, 16)
XY_POINTS = ((64, 0), (64, 38), (64, 63), (128, 0), (128, 63))   # (L, rung of linspace(0.7, 1.2, 64))
XY_SMALL_POINTS = tuple((L_, r) for L_ in (16, 32) for r in (0, 13, 26, 38, 51, 63))
TRI_POINTS = ((31, 0), (22, 0), (26, 10), (18, 16), (14, 31), (31, 24))  # (T idx, B idx) of the 32 x 32 grid
CUBIC_T_IDX = (51, 61, 76)                     # of linspace(1.2, 1.7, 128)

SECTIONS = {
    # name: (nthermal, nsweep, runs per point)
    "ising_square32": (20000, 50000, 16),
    "xy_square_large": (4000, 12000, 16),
    "tri_dmi128": (1500, 3000, 12),
    "cubic16": (3000, 20000, 16),
    "xy_square_small": (2000, 12000, 16),
}


def points(section):
    """[(key, T, h, lattice spec)] of a section."""
    if section == "ising_square32":
        Ts = np.linspace(1.5, 3.5, 16)
        return [("T%d" % i, float(Ts[i]), 0.0, ("square_ising", 32)) for i in ISING_T_IDX]
    if section in ("xy_square_large", "xy_square_small"):
        Ts = np.linspace(0.7, 1.2, 64)
        pts = XY_POINTS if section == "xy_square_large" else XY_SMALL_POINTS
        return [("L%d_r%d" % (L_, r), float(Ts[r]), 0.0, ("square_xy", L_)) for L_, r in pts]
    if section == "tri_dmi128":
        Ts, Bs = np.linspace(0.05, 1.0, 32), np.linspace(0.0, 1.0, 32)
        return [("T%d_B%d" % (i, j), float(Ts[i]), float(Bs[j]), ("tri_dmi", 128)) for i, j in TRI_POINTS]
    if section == "cubic16":
        Ts = np.linspace(1.2, 1.7, 128)
        return [("T%d" % i, float(Ts[i]), 0.0, ("cubic", 16)) for i in CUBIC_T_IDX]
    raise ValueError(section)


_CACHE = {}


def lattice(spec):
    if spec in _CACHE:
        return _CACHE[spec]
    from mcsolver_amd import lattice as lat
    from mcsolver_amd import workloads as W
    kind, L_ = spec
    if kind == "square_ising":
        Lt = lat.square(L_, J=-1.0, ising=True)
    elif kind == "square_xy":
        Lt = lat.square(L_, J=-1.0)
    elif kind == "tri_dmi":
        Lt = W.triangular_neel_dmi(L_)
    elif kind == "cubic":
        Lt = lat.cubic(L_, J=-1.0)
    else:
        raise ValueError(kind)
    if kind != "square_ising":
        Lt.groups = [[0]]    # one group, the whole (one-orbital) lattice: <M^2>, <M^4> of the total spin
    _CACHE[spec] = Lt
    return Lt


def one(job):
    section, key, T, h, spec, seed, rng = job
    from mcsolver_amd import lattice as lat
    from oracle import mcref
    nth, nsw, _ = SECTIONS[section]
    Lt = lattice(spec)
    with mcref.rng(rng):
        if spec[0] == "square_ising":
            out, _ = mcref.ising_run(lat.ising_args(Lt, T, h, nth, nsw), seed=seed)
            return [float(v) for v in out]
        nc = 2 if spec[0] == "square_xy" else 3
        out, _, g = mcref.on_run(nc, lat.on_args(Lt, T, h, nth, nsw), seed=seed)
    return [float(out[8]), float(out[9]), float(out[22]), float(out[26]), float(g[3]), float(g[5])]


def main(sections, rng):
    OUT = out_path(rng)
    doc = json.load(open(OUT)) if os.path.exists(OUT) else {}
    jobs = []
    for sec in sections:
        si = list(SECTIONS).index(sec)   # seeds depend on the section, not on the call
        R = SECTIONS[sec][2]
        for pi, (key, T, h, spec) in enumerate(points(sec)):
            for r in range(R):
                jobs.append((sec, key, T, h, spec, 500000 + 10000 * si + 100 * pi + r, rng))
    # longest first
    cost = {"ising_square32": 1, "xy_square_large": 4, "xy_square_small": 4, "tri_dmi128": 6, "cubic16": 4}
    jobs.sort(key=lambda j: -cost[j[0]] * (j[4][1] ** (3 if j[4][0] == "cubic" else 2)))
    t0 = time.time()
    with mp.get_context("fork").Pool(min(8, os.cpu_count() or 1)) as pool:
        res = pool.map(one, jobs, chunksize=1)
    for sec in sections:
        nth, nsw, R = SECTIONS[sec]
        pts = {}
        for (key, T, h, spec) in points(sec):
            pts[key] = {"T": T, "h": h, "runs": []}
        for job, v in zip(jobs, res):
            if job[0] == sec:
                pts[job[1]]["runs"].append(v)
        doc[sec] = {"nthermal": nth, "nsweep": nsw, "runs_per_point": R,
                    "values": (["spin_i", "spin_j", "spin_ij", "autoCorr", "E", "E2", "E_r", "E2_r", "U4",
                                "spin_tot"] if sec == "ising_square32" else ["E", "E2", "absM", "Q", "M2", "M4"]),
                    "points": pts}
    doc["generator"] = ("oracle/ref_glibc.c mcref_ising_run / mcref_on_run via oracle/mcref.py, FP64, integer "
                        "draws from %s" % ("xoshiro256** (top 31 bits)" if rng == "xoshiro" else "glibc rand()"))
    doc["rng"] = rng
    with open(OUT, "w") as f:
        json.dump(doc, f, indent=0)
    print("wrote %s (%d runs, %.0f s)" % (OUT, len(jobs), time.time() - t0))


if __name__ == "__main__":
    args = sys.argv[1:]
    rng = "xoshiro"
    if args and args[0].startswith("--rng="):
        rng = args.pop(0).split("=", 1)[1]
    main(args or list(SECTIONS), rng)
