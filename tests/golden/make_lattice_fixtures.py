"""Record the MCMainFunction argument tuples that the reference's own Python
builds (Lattice.py neighbour search + mcMain.MC flattening), as fixtures
for tests/test_lattice_fixtures.py.

Runs only where /root/reference exists (this container): it imports the
reference's mcsolver/Lattice.py and mcsolver/mcMain.py (and, for the sample
inputs, win.py and fileio.py) with
  * tkinter and matplotlib's Tk backend replaced by inert stand-ins (the GUI
    is not exercised; mcMain.py:8-13 imports win, which imports tkinter), and
  * recording stand-ins for the compiled isinglib / xylib / heisenberglib
    (the names mcMain.py:110-113 and 228-237 import first), which store the
    positional arguments of every MCMainFunction call and return a tuple of
    the reference's shape (isingLib.c:436-447, heisenbergLib.c:854-883).
No reference C is built or run here.

Two kinds of cases:
  tiny   lattices given directly to mcMain.MC (including L <= 2 max|overLat|,
         where Orbital.addLinking merges or skips repeated targets,
         Lattice.py:34-57); the full argument tuples are stored.
  sample the four input files under tests/golden/samples (copies of the
         reference's samples/*), run through the reference's own
         win.startSimulation(updateGUI=False) with a serial stand-in for
         multiprocessing.Pool; every call's arguments are stored as SHA-256
         digests plus lengths and scalars.

    python tests/golden/make_lattice_fixtures.py     # writes lattice_fixtures.json
"""
import hashlib
import json
import os
import shutil
import sys
import tempfile
import types
from unittest import mock

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/mcsolver"
OUT = os.path.join(HERE, "lattice_fixtures.json")
SAMPLES = os.path.join(HERE, "samples")


def digest(v):
    """(kind, length, sha256) of a tuple argument, or the scalar itself."""
    if isinstance(v, tuple):
        isf = any(isinstance(x, float) for x in v)
        a = np.asarray(v, np.float64 if isf else np.int64)
        return ["f" if isf else "i", len(v), hashlib.sha256(a.tobytes()).hexdigest()]
    if v is None or callable(v):
        return None
    return v.item() if isinstance(v, np.generic) else v


def plain(v):
    if isinstance(v, tuple):
        return [float(x) if isinstance(x, float) else int(x) for x in v]
    if v is None or callable(v):
        return None
    return v.item() if isinstance(v, np.generic) else v


# ---------------------------------------------------------------------------
# tiny lattices (specs shared with the test through the fixture file)
# ---------------------------------------------------------------------------
def _rng_tensor(rng, diag=-1.0):
    J = np.zeros(9)
    J[:3] = diag + rng.uniform(-0.1, 0.1, 3)
    J[3:] = rng.uniform(-0.05, 0.05, 6)
    return [float(x) for x in J]


def _dmi_bonds(D0=0.3):
    a1, a2 = np.array([1.0, 0, 0]), np.array([0.5, np.sqrt(3) / 2, 0])
    out = []
    for o in [(1, 0, 0), (0, 1, 0), (-1, 1, 0)]:
        r = o[0] * a1 + o[1] * a2
        r = r / np.linalg.norm(r)
        Dv = -D0 * np.cross([0, 0, 1.0], r)
        M = -np.eye(3) + np.array([[0, Dv[2], -Dv[1]], [-Dv[2], 0, Dv[0]], [Dv[1], -Dv[0], 0]])
        # reference order xx, yy, zz, xy, xz, yz, yx, zx, zy
        J = [M[0, 0], M[1, 1], M[2, 2], M[0, 1], M[0, 2], M[1, 2], M[1, 0], M[2, 0], M[2, 1]]
        out.append([0, 0, list(o), [float(x) for x in J]])
    return out


def tiny_cases():
    rng = np.random.default_rng(20261203)
    sq = [[1, 0, 0], [0, 1, 0], [0, 0, 1]]
    hexm = [[1, 0, 0], [-0.5, float(np.sqrt(3) / 2), 0], [0, 0, 1]]
    tri = [[1, 0, 0], [0.5, float(np.sqrt(3) / 2), 0], [0, 0, 1]]
    hc_bonds = [[s, t, list(o), _rng_tensor(rng)] for (s, t, o) in
                [(0, 1, (0, 0, 0)), (0, 1, (-1, 0, 0)), (0, 1, (0, 1, 0)), (0, 0, (1, 0, 0)), (0, 0, (1, 1, 0)),
                 (0, 0, (0, 1, 0)), (1, 1, (1, 0, 0)), (1, 1, (1, 1, 0)), (1, 1, (0, 1, 0))]]
    base = dict(ki_s=0, ki_t=0, ki_overLat=[0, 0, 0], orbGroupList=[], groupInSC=False, localCircuitList=[],
                nthermal=10, nsweep=20, ninterval=-1, algorithm="Metropolis", spinFrame=0, D=None)
    cases = [
        dict(name="ising_square4", model="Ising", L=[4, 4, 1], LMatrix=sq, pos=[[0, 0, 0]], S=[1.0],
             bonds=[[0, 0, [1, 0, 0], [-1.0]], [0, 0, [0, 1, 0], [-1.0]]], T=2.0, h=0.1),
        dict(name="ising_square2_merge", model="Ising", L=[2, 2, 1], LMatrix=sq, pos=[[0, 0, 0]], S=[1.0],
             bonds=[[0, 0, [1, 0, 0], [-1.0]], [0, 0, [0, 1, 0], [-1.0]], [0, 0, [-1, 0, 0], [-0.5]]], T=1.5, h=0.0,
             algorithm="Wolff"),
        dict(name="ising_square2_numpyT", model="Ising", L=[2, 2, 1], LMatrix=sq, pos=[[0, 0, 0]], S=[1.0],
             bonds=[[0, 0, [1, 0, 0], [-1.0]], [0, 0, [0, 1, 0], [-1.0]]], T=1.5, h=0.0, T_numpy=True),
        dict(name="ising_2orb3", model="Ising", L=[3, 3, 1], LMatrix=sq, pos=[[0, 0, 0], [0.5, 0, 0]], S=[1.0, 1.5],
             bonds=[[0, 1, [0, 0, 0], [-1.0]], [1, 0, [1, 0, 0], [-0.7]], [0, 0, [0, 1, 0], [0.3]],
                    [1, 1, [0, 1, 0], [-1.2]]], T=1.2, h=0.05, ki_s=0, ki_t=1, ki_overLat=[1, 0, 0]),
        dict(name="heis_honeycomb2_aniso", model="Heisenberg", L=[2, 2, 1], LMatrix=hexm,
             pos=[[1 / 3, 2 / 3, 0], [2 / 3, 1 / 3, 0]], S=[1.5, 1.5], D=[[0, 0, -0.1], [0.02, 0, -0.1]],
             bonds=hc_bonds, T=0.8, h=0.2, orbGroupList=[[0], [1]], groupInSC=True,
             localCircuitList=[[[0, [0, 0, 0]], [1, [0, 0, 0]], [0, [1, 0, 0]]]]),
        dict(name="heis_honeycomb3_aniso", model="Heisenberg", L=[3, 3, 1], LMatrix=hexm,
             pos=[[1 / 3, 2 / 3, 0], [2 / 3, 1 / 3, 0]], S=[1.5, 1.0], D=[[0, 0, -0.1], [0.0, 0.03, 0]],
             bonds=hc_bonds, T=1.3, h=0.0, orbGroupList=[[0, 1]], groupInSC=False, ki_s=0, ki_t=1,
             ki_overLat=[0, 1, 0]),
        dict(name="tri_dmi2", model="Heisenberg", L=[2, 2, 1], LMatrix=tri, pos=[[0, 0, 0]], S=[1.0],
             D=[[0, 0, -0.05]], bonds=_dmi_bonds(), T=0.5, h=0.4,
             localCircuitList=[[[0, [0, 0, 0]], [0, [1, 0, 0]], [0, [0, 1, 0]]],
                               [[0, [1, 0, 0]], [0, [1, 1, 0]], [0, [0, 1, 0]]]]),
        dict(name="tri_dmi3", model="Heisenberg", L=[3, 3, 1], LMatrix=tri, pos=[[0, 0, 0]], S=[1.0],
             D=[[0, 0, -0.05]], bonds=_dmi_bonds(), T=0.05, h=1.0,
             localCircuitList=[[[0, [0, 0, 0]], [0, [1, 0, 0]], [0, [0, 1, 0]]],
                               [[0, [1, 0, 0]], [0, [1, 1, 0]], [0, [0, 1, 0]]]]),
        dict(name="heis_cubic2", model="Heisenberg", L=[2, 2, 2], LMatrix=sq, pos=[[0, 0, 0]], S=[1.0],
             D=[[0, 0, 0]], bonds=[[0, 0, list(o), [-1.0, -1.0, -1.0, 0, 0, 0, 0, 0, 0]]
                                   for o in ([1, 0, 0], [0, 1, 0], [0, 0, 1])], T=1.44, h=0.0, algorithm="Wolff"),
        dict(name="xy_square3", model="XY", L=[3, 3, 1], LMatrix=sq, pos=[[0, 0, 0]], S=[1.0], D=[[0.05, 0, 0]],
             bonds=[[0, 0, [1, 0, 0], [-1.0, -1.0, -1.0, 0, 0, 0, 0, 0, 0]],
                    [0, 0, [0, 1, 0], [-1.0, -0.8, -1.0, 0.1, 0, 0, -0.1, 0, 0]]], T=0.9, h=0.1),
    ]
    out = []
    for c in cases:
        d = dict(base)
        d.update(c)
        if d["D"] is None:
            d["D"] = [[0.0, 0.0, 0.0]] * len(d["pos"])
        out.append(d)
    return out


# ---------------------------------------------------------------------------
# stand-ins
# ---------------------------------------------------------------------------
CALLS = []


def _fake_ising(*args):
    CALLS.append(("isinglib", args))
    n, sf = len(args[1]), args[14]
    frames = tuple(tuple(0.0 for _ in range(n)) for _ in range(sf)) if sf > 0 else 0.0
    return (0.0,) * 8 + (1.0, 0.0) + (frames,)


def _fake_on(name):
    def f(*args):
        CALLS.append((name, args))
        n, ng, sf = len(args[1]), args[12], args[20]
        frames = tuple(tuple((0.0, 0.0, 0.0) for _ in range(n)) for _ in range(sf)) if sf > 0 else 0.0
        groups = tuple(0.0 for _ in range((ng + 2) * (ng + 1))) if ng > 0 else 0.0
        return (0.0,) * 10 + (1.0,) + (0.0,) * 16 + (frames, groups)
    return f


class _SerialPool:
    """multiprocessing.Pool without processes: each task's parameters are
    pickled and unpickled as a worker receives them, so that every task
    scales its own copy of the bonds (mcMain.py:33-35 modifies them)."""

    def __init__(self, processes=None):
        pass

    def imap_unordered(self, fn, items):
        import pickle
        for x in items:
            yield fn(pickle.loads(pickle.dumps(x)))

    def close(self):
        pass


def install():
    import matplotlib
    matplotlib.use("Agg")
    for name in ("tkinter", "tkinter.filedialog", "tkinter.ttk", "tkinter.messagebox",
                 "matplotlib.backends.backend_tkagg"):
        sys.modules[name] = mock.MagicMock()
    for name, fn in (("isinglib", _fake_ising), ("xylib", _fake_on("xylib")), ("heisenberglib", _fake_on("heisenberglib"))):
        m = types.ModuleType(name)
        m.MCMainFunction = fn
        sys.modules[name] = m
    sys.path.insert(0, REF)


def run_tiny(case):
    import Lattice as lat_ref
    import mcMain
    On = case["model"] != "Ising"
    bonds = []
    for s, t, o, J in case["bonds"]:
        Jv = list(J) + [0.0] * (9 - len(J))
        bonds.append(lat_ref.Bond(s, t, np.array(o), *Jv, On=On))
    Lx, Ly, Lz = case["L"]
    T = np.float64(case["T"]) if case.get("T_numpy") else case["T"]   # loadMC passes np.linspace values
    mc = mcMain.MC(0, case["LMatrix"], pos=case["pos"], S=case["S"], D=case["D"], bondList=bonds, T=T,
                   Lx=Lx, Ly=Ly, Lz=Lz, ki_s=case["ki_s"], ki_t=case["ki_t"], ki_overLat=case["ki_overLat"],
                   orbGroupList=case["orbGroupList"], groupInSC=case["groupInSC"], h=case["h"],
                   On={"Ising": 1, "XY": 2, "Heisenberg": 3}[case["model"]], spinFrame=case["spinFrame"],
                   localCircuitList=[tuple((o, tuple(r)) for o, r in c) for c in case["localCircuitList"]])
    n0 = len(CALLS)
    if case["model"] == "Ising":
        mc.mainLoopViaCLib(nsweep=case["nsweep"], nthermal=case["nthermal"], ninterval=case["ninterval"],
                           algo=case["algorithm"])
    else:
        mc.mainLoopViaCLib_On(nsweep=case["nsweep"], nthermal=case["nthermal"], ninterval=case["ninterval"],
                              algo=case["algorithm"], On=2 if case["model"] == "XY" else 3)
    assert len(CALLS) == n0 + 1
    return CALLS[-1]


def run_sample(path):
    import win
    win.Pool = _SerialPool
    n0 = len(CALLS)
    win.startSimulation(updateGUI=False, rpath=path)
    return CALLS[n0:]


def main():
    install()
    doc = {"tiny": [], "samples": {}}
    cwd = os.getcwd()
    work = tempfile.mkdtemp(prefix="mcsolver_ref_")
    try:
        os.chdir(work)  # mcMain and win write ./out, ./spinDotSpin.txt, ./result.txt
        for case in tiny_cases():
            lib, args = run_tiny(case)
            doc["tiny"].append({"spec": case, "lib": lib, "args": [plain(a) for a in args]})
            print("tiny", case["name"], lib, "n =", len(args[1]), flush=True)
        for name in sorted(os.listdir(SAMPLES)):
            shutil.copy(os.path.join(SAMPLES, name), os.path.join(work, name))
            calls = run_sample(os.path.join(work, name))
            doc["samples"][name] = [{"lib": lib, "args": [digest(a) for a in args]} for lib, args in calls]
            print("sample", name, len(calls), "calls", flush=True)
    finally:
        os.chdir(cwd)
        shutil.rmtree(work, ignore_errors=True)
    doc["generator"] = ("reference mcsolver/Lattice.py + mcMain.py (+ win.py, fileio.py for the samples) with "
                        "recording stand-ins for the compiled libraries; tkinter stubbed")
    with open(OUT, "w") as f:
        json.dump(doc, f, separators=(",", ":"))
    print("wrote", OUT)


if __name__ == "__main__":
    main()
