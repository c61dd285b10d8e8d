"""The FP64 honeycomb fixture (tests/golden/honeycomb_fp64.json) is what its
generator writes: one of its runs regenerates bit for bit, and every run obeys
the bounds of the model (two unit spins per cell)."""
import json
import os

import numpy as np

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "honeycomb_fp64.json")


def test_fixture_bounds():
    d = json.load(open(FIX))
    assert len(d["temps"]) == 4 and d["runs_per_T"] >= 6
    for T in d["temps"]:
        E, E2, M, M2, M4 = np.array(d["runs"]["%.6f" % T]).T
        assert (E < 0).all() and (E2 >= E * E).all()
        assert (M > 0).all() and (M <= 2.0 + 1e-12).all()
        assert (M2 >= M * M - 1e-12).all() and (M4 >= M2 * M2 - 1e-12).all()


def test_fixture_run_regenerates():
    import make_honeycomb_fp64 as g
    d = json.load(open(FIX))
    T = d["temps"][0]
    assert (g.NTHERMAL, g.NSWEEP) == (d["nthermal"], d["nsweep"])
    got = g.one((T, g.SEED0, d["rng"]))
    assert got == d["runs"]["%.6f" % T][0]
