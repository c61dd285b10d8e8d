"""The FP64 oracle restatement replays the reference's own compiled C engines.

Fixtures (tests/golden/replay_fixtures.json) were written by
tests/golden/make_fixtures.py from mcsolver/isingLib.c and
mcsolver/heisenbergLib.c compiled unmodified; the same glibc srand() seed
drives both, so the outputs must agree to the last bit (tolerance 1e-12 is
only there to survive a different libm).
"""
import json
import os

import numpy as np
import pytest

from cases import CASES, args_digest, case_args
from oracle import mcref

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "replay_fixtures.json")))


def _flat(x):
    if isinstance(x, (list, tuple)):
        out = []
        for v in x:
            out += _flat(v)
        return out
    return [float(x)]


@pytest.mark.parametrize("name", sorted(CASES))
def test_replay_matches_reference(name):
    fx = FIX[name]
    args = case_args(name)
    # the builder must produce the very inputs the reference consumed
    assert args_digest(args) == fx["args_sha256"]
    if fx["model"] == "ising":
        ours = mcref.ising_tuple(args, seed=fx["seed"])
    else:
        ours = mcref.on_tuple(3, args, seed=fx["seed"])
    ref = fx["result"]
    assert len(ours) == len(ref)
    a, b = np.array(_flat(list(ours))), np.array(_flat(ref))
    assert a.shape == b.shape
    np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12)


def test_fixture_cases_exercise_quirks():
    # the linkStrength+i quirk matters only with unequal couplings
    args = case_args("ising_aniso_quirk")
    J = np.array(args[7]).reshape(-1, args[5])
    assert len(np.unique(J[J != 0])) > 1
    # ninterval < N exercises the partial totSpin start (isingLib.c:350)
    assert case_args("ising_2orb_ninterval")[4] < len(case_args("ising_2orb_ninterval")[1])
