"""The lattice builder and the tuple flattening against the reference's own
Python: tests/golden/lattice_fixtures.json holds the MCMainFunction
arguments that mcsolver/Lattice.py + mcMain.MC built (recorded by
tests/golden/make_lattice_fixtures.py with stand-ins for the compiled
libraries).  lattice.build_lattice + lattice.ising_args / on_args (and, for
the sample input files, driver.load_input + driver.build_input_lattice)
must reproduce them exactly: the neighbour order of Lattice.py:237-284, the
merge / skip of repeated targets of Lattice.py:34-57 (including the in-place
addition into the bond's shared tensor), the -1 / 0.0 padding, the
block-spin tables, circuits, groups and the beta scaling of mcMain.py:16-248.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from mcsolver_amd import driver as D
from mcsolver_amd import lattice as lat

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "lattice_fixtures.json")
SAMPLES = os.path.join(HERE, "golden", "samples")
_DOC = json.load(open(FIX))
TINY = {c["spec"]["name"]: c for c in _DOC["tiny"]}

ARG_NAMES = {
    "isinglib": ["algorithm", "initSpin", "nthermal", "nsweep", "ninterval", "maxNLinking", "nlink", "linkStrength",
                 "linkedOrb", "corrOrbPair", "h", "rOrb", "rOrbCluster", "linkedOrb_rnorm", "spinFrame", "callback"],
    "on": ["algorithm", "initSpin", "initD", "nthermal", "nsweep", "ninterval", "maxNLinking", "nlink",
           "linkStrength", "linkedOrb", "localCircuits", "corrOrbPair", "nOrbGroup", "maxOrbGroupSize",
           "orbGroupList", "flunc", "h", "rOrb", "rOrbCluster", "linkedOrb_rnorm", "spinFrame",
           "ignoreNonDiagonalJ", "callback"],
}


def _ours(spec):
    ising = spec["model"] == "Ising"
    bonds = [lat.Bond(s, t, tuple(o), J) for s, t, o, J in spec["bonds"]]
    Lt = lat.build_lattice(*spec["L"], spec["LMatrix"], spec["pos"], bonds, spec["S"], spec["D"],
                           ki_s=spec["ki_s"], ki_t=spec["ki_t"], ki_overLat=tuple(spec["ki_overLat"]),
                           orbGroupList=spec["orbGroupList"], groupInSC=spec["groupInSC"],
                           localCircuitList=[tuple((o, tuple(r)) for o, r in c) for c in spec["localCircuitList"]],
                           ising=ising)
    f = lat.ising_args if ising else lat.on_args
    T = np.float64(spec["T"]) if spec.get("T_numpy") else spec["T"]
    return f(Lt, T, spec["h"], spec["nthermal"], spec["nsweep"], spec["ninterval"], spec["algorithm"],
             spec["spinFrame"])


def _same(name, got, ref):
    if isinstance(ref, list):
        g = list(got)
        assert len(g) == len(ref), "%s: length %d, reference %d" % (name, len(g), len(ref))
        bad = [k for k, (a, b) in enumerate(zip(g, ref)) if a != b]
        assert not bad, "%s differs at %d entries, first %d: ours %r reference %r" % (
            name, len(bad), bad[0], g[bad[0]], ref[bad[0]])
    else:
        assert got == ref or (ref is None and got is None), "%s: ours %r reference %r" % (name, got, ref)


@pytest.mark.parametrize("name", sorted(TINY))
def test_tiny_lattice_tuples_match_reference(name):
    c = TINY[name]
    names = ARG_NAMES["isinglib" if c["lib"] == "isinglib" else "on"]
    got = _ours(c["spec"])
    assert len(got) == len(c["args"]) == len(names)
    for nm, g, r in zip(names, got, c["args"]):
        if nm == "callback":
            continue
        _same("%s.%s" % (name, nm), g, r)


def _digest(v):
    if isinstance(v, tuple):
        isf = any(isinstance(x, float) for x in v)
        a = np.asarray(v, np.float64 if isf else np.int64)
        return ["f" if isf else "i", len(v), hashlib.sha256(a.tobytes()).hexdigest()]
    if v is None or callable(v):
        return None
    return v


@pytest.mark.parametrize("sample", sorted(_DOC["samples"]))
def test_sample_input_tuples_match_reference(sample):
    """Every (H, T) point of a sample input, in win.startSimulation's order
    (H outer, T inner; win.py:69-72, 113-116)."""
    calls = _DOC["samples"][sample]
    p = D.load_input(os.path.join(SAMPLES, sample))
    Lt = D.build_input_lattice(p)
    ising = p.modelType == "Ising"
    names = ARG_NAMES["isinglib" if ising else "on"]
    k = 0
    for H in np.linspace(p.H0, p.H1, p.nH):
        for T in np.linspace(p.T0, p.T1, p.nT):
            f = lat.ising_args if ising else lat.on_args
            got = f(Lt, T, H, p.nthermal, p.nsweep, p.ninterval, p.algorithm, p.spinFrame)
            rec = calls[k]
            for nm, g, r in zip(names, got, rec["args"]):
                if nm == "callback":
                    continue
                assert _digest(g) == r, "%s call %d (T=%g, H=%g): %s differs" % (sample, k, T, H, nm)
            k += 1
    assert k == len(calls)
