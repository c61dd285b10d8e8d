"""FP64 oracle estimates on the CrI3-like honeycomb lattice (64x64 cells,
full-tensor first- and second-neighbour bonds, Dz = -0.1): the reference for
the FP32 / FP16-storage bias report in tests/test_gpu_fp16_bias.py.

For each temperature, R independent runs of the oracle restatement of
heisenberglib.MCMainFunction (oracle/ref_glibc.c, random-sequential
Metropolis with glibc rand(), FP64) record the per-run moments
  E = tuple[8], E2 = tuple[9], |M| = tuple[22] (total spin along its own
  direction / nLat), M2 and M4 of the total-spin group (groups output),
from which the test forms E, C, |M|, chi and U4.  The oracle is our own C
code, so this runs anywhere the oracle builds; one process per run.

Integer draws from xoshiro256** (default, honeycomb_fp64.json) or, with
--rng=glibc, from glibc rand() as the reference (honeycomb_fp64_glibc.json);
see make_config_fp64.py for why the parity reference uses the former.

    python tests/golden/make_honeycomb_fp64.py [--rng=xoshiro|glibc]
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
    return os.path.join(HERE, "honeycomb_fp64.json" if rng == "xoshiro" else "honeycomb_fp64_%s.json" % rng)
TEMPS_IDX = (10, 30, 45, 60)      # of linspace(0.1, 2.0, 64), the honeycomb workload's ladder
NTHERMAL, NSWEEP, R = 500, 2000, 32
SEED0 = 31000


def lattice():
    from mcsolver_amd import workloads as W
    H = W.honeycomb_full_tensor(64)
    H.groups = [[0, 1]]           # one orbital group: the groups output then carries <M^2>, <M^4> of the total
    return H


def one(job):
    T, seed, rng = job
    from mcsolver_amd import lattice as lat
    from oracle import mcref
    H = lattice()
    a = lat.on_args(H, T, 0.0, NTHERMAL, NSWEEP)
    with mcref.rng(rng):
        out, _, g = mcref.on_run(3, a, seed=seed)
    return [float(out[8]), float(out[9]), float(out[22]), float(g[3]), float(g[5])]


def main(rng="xoshiro"):
    temps = [float(t) for t in np.linspace(0.1, 2.0, 64)[list(TEMPS_IDX)]]
    jobs = [(T, SEED0 + 100 * i + r, rng) for i, T in enumerate(temps) for r in range(R)]
    t0 = time.time()
    with mp.get_context("fork").Pool(min(8, os.cpu_count() or 1)) as pool:
        res = pool.map(one, jobs)
    runs = {}
    for (T, _, _), v in zip(jobs, res):
        runs.setdefault("%.6f" % T, []).append(v)
    H = lattice()
    doc = {"lattice": "honeycomb 64x64 cells (8192 sites), full 3x3 tensors, Dz=-0.1, seed 0x5EED2026",
           "nLat": int(H.n // 2), "n": int(H.n), "nthermal": NTHERMAL, "nsweep": NSWEEP, "runs_per_T": R,
           "moments": ["E", "E2", "absM", "M2", "M4"], "temps": temps, "runs": runs,
           "generator": "oracle/ref_glibc.c mcref_on_run via oracle/mcref.py, FP64, integer draws from %s" %
                        ("xoshiro256** (top 31 bits)" if rng == "xoshiro" else "glibc rand()"), "rng": rng}
    OUT = out_path(rng)
    with open(OUT, "w") as f:
        json.dump(doc, f, indent=1)
    print("wrote %s (%d runs, %.0f s)" % (OUT, len(jobs), time.time() - t0))


if __name__ == "__main__":
    main(sys.argv[1].split("=", 1)[1] if len(sys.argv) > 1 and sys.argv[1].startswith("--rng=") else "xoshiro")
