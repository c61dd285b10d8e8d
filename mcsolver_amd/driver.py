"""Batched replacement for ``mcsolver.loadMC`` and ``mcMain.MC``.

The reference runs one process per (T, H) point (win.py:40-159): each builds
the lattice, scales every energy by 1/T (mcMain.py:16-49), calls
MCMainFunction once, and appends its line to ``./out`` (mcMain.py:130-132,
263-265), writes spin-distribution and group files (mcMain.py:274-323); at
the end ``./result.txt`` gets one row per point (win.py:151-156).

Here the lattice is built once with unscaled couplings.  Every (T, H) point
becomes one replica of one Batch, with its own row beta = 1/T and field H,
and all points run in the same engine launches.  The engine's accumulators
are turned into exactly the tuples MCMainFunction returns (energies in
beta units, as the reference has them before mcMain multiplies by T), and
the post-processing and file formats follow mcMain/win line for line.

Differences from the reference, all deliberate:
* result.txt rows are in point order (iH outer, iT inner); the reference
  writes them in Pool completion order, which equals this order only with
  one core;
* caller-supplied bond objects are not rescaled in place (the reference's
  Bond.renormWithT mutates them, Lattice.py:140-147);
* a dipolar coupling alpha != 0 adds truncated dipolar bond entries
  (lattice.dipolar_bonds); the reference's generateDipoleBondings raises
  TypeError for Ising and exits for O(n) (Lattice.py:286-308);
* input errors raise ValueError instead of printing and returning.
"""
from __future__ import annotations

import os
import re
import time
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import _lib as L
from . import lattice as lat
from .dropin import (DEFAULT_SEED, _plan_counts, _run_frames, check_health, recompute_period, ising_couplings, ising_meas_flags, ising_observables,
                     on_initial_spins, on_meas_flags, on_observables)
from .engine import Batch, ReplicaSpec, compile_graph

SETTING_FILE_VERSION = "3.0"   # win.py:19
MODELS = {"Ising": 1, "XY": 2, "Heisenberg": 3}

_NUM = re.compile(r"[0-9\.\-]+")
_UINT = re.compile(r"[0-9]+")
_INT = re.compile(r"[0-9\-]+")
_UFLOAT = re.compile(r"[0-9\.]+")
_WORD = re.compile(r"[a-zA-Z]+")


# ---------------------------------------------------------------------------
# input file (fileio.py:169-292 semantics)
# ---------------------------------------------------------------------------

@dataclass
class McInput:
    """Everything a v3.0 input file holds (fileio.py:169-292)."""

    LMatrix: List[List[float]]
    LPack: List[int]
    pos: List[List[float]]
    S: List[float]
    DList: List[List[float]]
    bondList: List[list]        # [source, target, [ox, oy, oz], Jxx, Jyy, Jzz, Jxy, Jxz, Jyz, Jyx, Jzx, Jzy]
    T0: float
    T1: float
    nT: int
    H0: float
    H1: float
    nH: int
    dipoleAlpha: float
    GcOrb: list                 # [[ki_s, ki_t], [ox, oy, oz]]
    orbGroupList: List[List[int]]
    groupInSC: bool
    localCircuitList: list      # [((orb, (ox, oy, oz)), x3), ...]
    spinFrame: int
    nthermal: int
    nsweep: int
    ninterval: int
    xAxisType: str
    modelType: str
    algorithm: str
    ncores: int = 1


_TAGS = ("Lattice", "Supercell", "Orbitals", "Bonds", "Temperature", "Field", "Dipole", "Sweeps", "Model",
         "Algorithm", "Measurement", "OrbGroup", "Ncores", "Distribution", "XAxis", "LocalCircuit")
_REQUIRED = ("Lattice", "Supercell", "Orbitals", "Bonds", "Temperature", "Sweeps", "Model", "Algorithm", "Ncores",
             "Field", "Dipole", "Distribution", "XAxis", "OrbGroup", "LocalCircuit")


def parse_input(text: str) -> McInput:
    """Parse the text of a v3.0 input file.

    Same rules as fileio.loadParam (fileio.py:169-292): blank lines dropped;
    a section starts at the last line whose first word is its tag; values are
    read positionally with the same number patterns, so the stale column
    labels of some samples ("Jz Jx Jy", "Dz Dx Dy") do not matter: the nine
    bond numbers are always Jxx Jyy Jzz Jxy Jxz Jyz Jyx Jzx Jzy and the three
    orbital numbers after the position are always Dx Dy Dz.
    """
    data = [line for line in text.split("\n") if line]
    if not data:
        raise ValueError("empty input file")
    ver = _UFLOAT.findall(data[0])
    if not ver or ver[0] != SETTING_FILE_VERSION:
        raise ValueError("unknown file or version (only support v%s)" % SETTING_FILE_VERSION)
    tag = {}
    for i, line in enumerate(data):
        w = _WORD.findall(line)
        if w and w[0] in _TAGS:
            tag[w[0]] = i
    missing = [t for t in _REQUIRED if not tag.get(t)]
    if missing:
        raise ValueError("cannot find some tags: %s" % ", ".join(missing))

    def nums(i, pat=_NUM):
        return pat.findall(data[i])

    LMatrix = [[float(x) for x in nums(tag["Lattice"] + 1 + i)] for i in range(3)]
    LPack = [int(x) for x in nums(tag["Supercell"] + 1, _UINT)]
    norb = int(nums(tag["Orbitals"] + 1, _UINT)[0])
    pos, S, DList = [], [], []
    for i in range(norb):
        e = nums(tag["Orbitals"] + 3 + i)
        pos.append([float(e[3]), float(e[4]), float(e[5])])
        S.append(float(e[2]))
        DList.append([float(e[6]), float(e[7]), float(e[8])])
    nbonds = int(nums(tag["Bonds"] + 1, _UINT)[0])
    bondList = []
    for i in range(nbonds):
        e = nums(tag["Bonds"] + 3 + i)
        bondList.append([int(e[10]), int(e[11]), [int(e[12]), int(e[13]), int(e[14])]] +
                        [float(e[k]) for k in range(1, 10)])
    s, t, v1, v2, v3 = [int(x) for x in nums(tag["Measurement"] + 1, _INT)]
    GcOrb = [[s, t], [v1, v2, v3]]
    ng = int(nums(tag["OrbGroup"], _UINT)[0])
    groupInSC = data[tag["OrbGroup"] + 1] == "Supergroup"
    orbGroupList = []
    for i in range(ng):
        _, o0, o1 = nums(tag["OrbGroup"] + i + 2, _UINT)[:3]
        orbGroupList.append(list(range(int(o0), int(o1) + 1)))
    ncirc = int(nums(tag["LocalCircuit"], _UINT)[0])
    circuits = []
    for i in range(ncirc):
        e = [int(x) for x in nums(tag["LocalCircuit"] + 1 + i, _INT)]
        circuits.append(((e[1], (e[2], e[3], e[4])), (e[5], (e[6], e[7], e[8])), (e[9], (e[10], e[11], e[12]))))
    Tp = nums(tag["Temperature"] + 1, _UFLOAT)
    Hp = nums(tag["Field"] + 1)
    sw = nums(tag["Sweeps"] + 1)
    return McInput(LMatrix=LMatrix, LPack=LPack, pos=pos, S=S, DList=DList, bondList=bondList,
                   T0=float(Tp[0]), T1=float(Tp[1]), nT=int(Tp[2]), H0=float(Hp[0]), H1=float(Hp[1]),
                   nH=int(Hp[2]), dipoleAlpha=float(nums(tag["Dipole"] + 1)[0]), GcOrb=GcOrb,
                   orbGroupList=orbGroupList, groupInSC=groupInSC, localCircuitList=circuits,
                   spinFrame=int(nums(tag["Distribution"], _UINT)[0]), nthermal=int(sw[0]), nsweep=int(sw[1]),
                   ninterval=int(sw[2]), xAxisType=data[tag["XAxis"] + 1], modelType=data[tag["Model"] + 1],
                   algorithm=data[tag["Algorithm"] + 1], ncores=int(data[tag["Ncores"] + 1]))


def load_input(path: str) -> McInput:
    with open(path, "r") as f:
        return parse_input(f.read())


def format_input(p: McInput) -> str:
    """Text of a v3.0 input file in the layout fileio.saveParam writes
    (fileio.py:108-167), so that parse_input(format_input(p)) == p."""
    o = ["This is mcsolver's save file, version: 3.0", "Lattice:"]
    o += ["%.9f %.9f %.9f" % tuple(a) for a in p.LMatrix]
    o += ["Supercell used in MC simulations:", "%d %d %d" % tuple(p.LPack), "Orbitals in cell:", "%d" % len(p.pos),
          "Positions, initial spin states and onsite-anisotropy of every orbital:"]
    for i, (ps, s, d) in enumerate(zip(p.pos, p.S, p.DList)):
        o.append("orb %d: type %d spin %.9f pos [%.9f %.9f %.9f] Dx %.9f Dy %.9f Dz %.9f h %.9f"
                 % (i, 0, s, ps[0], ps[1], ps[2], d[0], d[1], d[2], 0.0))
    o += ["Bonds:", "%d" % len(p.bondList), "id, source, target, overLat, exchange matrix elements of each bond:"]
    for i, b in enumerate(p.bondList):
        o.append("bond %d: Jx %.9f Jy %.9f Jz %.9f Jxy %.9f Jxz %.9f Jyz %.9f Jyx %.9f Jzx %.9f Jzy %.9f "
                 "orb %d to orb %d over [%d %d %d]" % ((i,) + tuple(b[3:12]) + (b[0], b[1]) + tuple(b[2])))
    o += ["Temperature scanning region:", "Tmin %.9f Tmax %.9f nT %d" % (p.T0, p.T1, p.nT),
          "Field scanning region (in unit 1.48872 T, only if Kelvin and uB is used for energy and spin):",
          "Hmin %.9f Hmax %.9f nH %d" % (p.H0, p.H1, p.nH), "Dipole long-range coupling:",
          "alpha %.6f" % p.dipoleAlpha, "Measurement:",
          "measure the correlation function between orb%d and orb%d over [%d %d %d]"
          % (p.GcOrb[0][0], p.GcOrb[0][1], p.GcOrb[1][0], p.GcOrb[1][1], p.GcOrb[1][2]),
          "Supergroup", "OrbGroup:%d" % len(p.orbGroupList), "Supergroup" if p.groupInSC else "Unitcell"]
    for i, g in enumerate(p.orbGroupList):
        o.append("group%d orb%d-orb%d" % (i, g[0], g[-1]))
    o += [">>>       Topological section      <<<",
          "LocalCircuit per cell: %d (set to 0 to skip the calc. for topo. Q)" % len(p.localCircuitList)]
    for i, c in enumerate(p.localCircuitList):
        o.append("Circuit %d enclosed by orb %d [%d %d %d], orb %d [%d %d %d], and orb %d [%d %d %d]"
                 % (i, c[0][0], *c[0][1], c[1][0], *c[1][1], c[2][0], *c[2][1]))
    o += [">>>   End of Topological section   <<<", "Distribution output frame: %d" % p.spinFrame,
          "Sweeps for thermalization and statistics, and relaxiation step for each sweep:",
          "%d %d %d" % (p.nthermal, p.nsweep, p.ninterval), "XAxis type:", p.xAxisType, "Model type:", p.modelType,
          "Algorithm:", p.algorithm, "Ncores:", "%d" % p.ncores]
    return "\n".join(o) + "\n"


# ---------------------------------------------------------------------------
# lattice of an input
# ---------------------------------------------------------------------------

def _bonds(bondList, ising: bool) -> List[lat.Bond]:
    out = []
    for b in bondList:
        J = [float(b[3])] if ising else [float(x) for x in b[3:12]]
        out.append(lat.Bond(int(b[0]), int(b[1]), tuple(int(x) for x in b[2]), J))
    return out


def build_input_lattice(p: McInput, dipole_cutoff: Optional[float] = None) -> lat.Lattice:
    """The lattice win.startSimulation + MC.__init__ build (win.py:56-64,
    mcMain.py:27-40), with unscaled couplings."""
    ising = p.modelType == "Ising"
    Lt = lat.build_lattice(p.LPack[0], p.LPack[1], p.LPack[2], p.LMatrix, p.pos, _bonds(p.bondList, ising), p.S,
                           p.DList, ki_s=p.GcOrb[0][0], ki_t=p.GcOrb[0][1], ki_overLat=tuple(p.GcOrb[1]),
                           orbGroupList=p.orbGroupList, groupInSC=p.groupInSC,
                           localCircuitList=p.localCircuitList, ising=ising, numpy_T=True)
    if abs(p.dipoleAlpha) > 1e-5:
        Lt = lat.dipolar_bonds(Lt, p.dipoleAlpha, cutoff=dipole_cutoff, ising=ising)
    return Lt


# ---------------------------------------------------------------------------
# the batched run
# ---------------------------------------------------------------------------

@dataclass
class PointResult:
    """One (T, H) point: the tuple MCMainFunction would return (energies in
    beta units) plus the point's parameters."""

    ID: int
    T: float          # as requested
    H: float
    Tc: float         # clamped to >= 0.1 (mcMain.py:21)
    data: tuple


def _diag_flag(Lt: lat.Lattice, Tc: float) -> int:
    """ignoreNonDiagonalJ of mcMain.py:158-174 for the beta-scaled table."""
    valid = Lt.nbr >= 0
    Js = (1 / Tc) * Lt.J[valid]
    return 0 if np.any(np.abs(Js[:, 3:]) > 1e-6) else 1


def run_points(Lt: lat.Lattice, nc: int, algorithm: str, points: Sequence[tuple], nthermal: int, nsweep: int,
               ninterval: int, spinFrame: int = 0, seed: int = DEFAULT_SEED, store: str = "f32",
               tempering: bool = False, flunc: float = 0.0) -> List[PointResult]:
    """Run every (T, H) point of one lattice in one Batch and return the
    MCMainFunction tuple of each, in point order.

    points: (T, H) pairs.  tempering=True adds replica-exchange swaps every
    sweep between points of equal H (Metropolis only; the reference has no
    tempering, so this changes the sampling but not the observables'
    expectation)."""
    if algorithm not in ("Metropolis", "Wolff"):
        raise ValueError("only Metropolis and Wolff are supported (win.py:69, 109)")
    wolff = algorithm == "Wolff"
    n = Lt.n
    ninterval = n if ninterval <= 0 else int(ninterval)  # mcMain.py:54, 145
    Tc = [lat.clamp_T(float(T)) for T, _ in points]
    nr = len(Lt.rOrb)
    pairs = Lt.corr_pairs
    if nc == 1:
        Jeff = ising_couplings(np.where(Lt.nbr >= 0, Lt.J[:, :, 0], 0.0).reshape(-1), n, Lt.maxl, Lt.nlink)
        graphs = [compile_graph(1, Lt.nbr, Jeff[:, :, None], Lt.S, np.zeros((n, 3)), pairs=pairs, rOrb=Lt.rOrb,
                                rOrbCluster=Lt.rOrbCluster if nr else None, rlink=Lt.rlink if nr else None)]
        gidx = [0] * len(points)
        meas = ising_meas_flags(nr > 0, wolff)
        init = [np.ones(n)] * len(points)
    else:
        flags = [_diag_flag(Lt, t) for t in Tc]
        uniq = sorted(set(flags))
        circ = Lt.circuits if (nc == 3 and len(Lt.circuits)) else None
        graphs = [compile_graph(nc, Lt.nbr, Lt.J, Lt.S, Lt.D, pairs=pairs, circuits=circ, rOrb=Lt.rOrb,
                                rOrbCluster=Lt.rOrbCluster if nr else None, rlink=Lt.rlink if nr else None,
                                groups=Lt.groups, diag_only=f > 0) for f in uniq]
        gidx = [uniq.index(f) for f in flags]
        meas = on_meas_flags(nc, circ is not None, nr > 0, len(Lt.groups) > 0)
        init = [on_initial_spins(Lt.S, nc, flunc, seed)] * len(points)
    reps = []
    if tempering:
        if wolff:
            raise ValueError("tempering is implemented for Metropolis only")
        byH = {}
        for k, (_, H) in enumerate(points):
            byH.setdefault(float(H), []).append(k)
        ladders = list(byH.values())
        order = [k for lad in ladders for k in lad]
        if order != list(range(len(points))):
            raise ValueError("tempering needs the points of one field to be consecutive")
        row0 = 0
        for il, lad in enumerate(ladders):
            if len({gidx[k] for k in lad}) != 1:
                raise ValueError("tempering ladder mixes diagonal-only and full couplings")
            for r, k in enumerate(lad):
                reps.append(ReplicaSpec(gidx[k], row0, nrung=len(lad), ladder=il, rung=r))
            row0 += len(lad)
    else:
        reps = [ReplicaSpec(gidx[k], k) for k in range(len(points))]
    b = Batch(nc, graphs, reps, [1.0 / t for t in Tc], [float(H) for _, H in points], init, store=store, meas=meas,
              nbins=1, nframes=int(nsweep), seed=seed, wolff=wolff,
              recompute_every=recompute_period(Lt.nbr, Lt.J, nc == 1))
    b.init()
    nth, spf, ncl = _plan_counts(n, int(nthermal), int(nsweep), ninterval, wolff)
    alg = L.WOLFF if wolff else L.METROPOLIS
    if tempering:
        frames = _run_frames_pt(b, nth, int(nsweep), spf, int(spinFrame), len(points))
    else:
        frames = _run_frames(b, nth, int(nsweep), spf, alg, ncl, int(spinFrame), replicas=range(len(points)))
    A = b.accumulators()
    G = b.group_acc() if (nc > 1 and len(Lt.groups)) else None
    if tempering:
        # accumulators are per row; row order == point order here
        pass
    check_health(b)
    out = []
    for k, (T, H) in enumerate(points):
        t = Tc[k]
        a = A[k].copy()
        # engine energies are in coupling units; MCMainFunction's are beta-scaled
        for q, p in ((L.A_E, 1), (L.A_E2, 2), (L.A_ER, 1), (L.A_ER2, 2)):
            a[q] /= t ** p
        if nc == 1:
            obs = ising_observables(a, int(nsweep), Lt.S, ninterval, wolff, len(pairs))
            fr = tuple(tuple(float(v) for v in (f[:, 0] * Lt.S)) for f in frames[k]) if spinFrame > 0 else 0.0
            data = tuple(float(v) for v in obs) + (fr,)
        else:
            vals = on_observables(a, int(nsweep), nc, nr > 0)
            if spinFrame > 0:
                fr = tuple(tuple((float(v[0]), float(v[1]), float(v[2]) if nc == 3 else 0.0) for v in f)
                           for f in frames[k])
            else:
                fr = 0.0
            gr = tuple(float(v) / float(nsweep) for v in G[k]) if G is not None else 0.0
            data = tuple(vals) + (fr, gr)
        out.append(PointResult(k, float(T), float(H), t, data))
    return out


def _run_frames_pt(b: Batch, nth: int, nsweep: int, spf: int, spinFrame: int, nrow: int) -> list:
    """Tempering run; spin frames are taken per temperature row (the
    replica currently at that row)."""
    frames = [[] for _ in range(nrow)]
    if spinFrame <= 0:
        b.run(nth, nsweep, spf, pt_every=1)
        return frames
    ops = nsweep // spinFrame
    if ops <= 0:
        raise ValueError("spinFrame > nsweep (the reference divides by zero here)")
    marks = list(range(0, nsweep, ops))
    done, sweep0, first = 0, 0, True
    for m in marks + [nsweep]:
        todo = m + 1 - done if m < nsweep else nsweep - done
        if todo > 0:
            th = nth if first else 0
            b.run(th, todo, spf, pt_every=1, sweep0=sweep0, frame0=done)
            sweep0 += th + todo * spf
            done += todo
            first = False
        if m < nsweep and len(frames[0]) < spinFrame:
            who = b.who.cpu().numpy()
            for row in range(nrow):
                frames[row].append(b.replica_spins(int(who[row])).copy())
    return frames


# ---------------------------------------------------------------------------
# output files (mcMain.py:126-135, 250-323; win.py:151-156)
# ---------------------------------------------------------------------------

class _Writer:
    def __init__(self, outdir: str, verbose: bool):
        self.d = outdir
        self.verbose = verbose

    def path(self, name):
        return os.path.join(self.d, name)

    def headers(self):
        for name in ("out", "spinDotSpin.txt"):
            with open(self.path(name), "w") as f:
                f.write("#T #H\n")

    def ising(self, Lt, r: PointResult, spinFrame: int):
        si, sj, sij, auto, E, E2, Er, E2r, U4, stot, frames = r.data
        T, h = r.Tc, r.H
        E *= T; E2 *= T ** 2; Er *= T; E2r *= T ** 2
        if self.verbose:
            print("%.3f %.3f %.3f %.3f %.3f %.3f %.6f %.3f %.3f %.3f %.3f %.6f %.6f %.6f" % (
                T, h, si, sj, stot, sij, auto, E, E2, Er, E2r, E2 - E ** 2, E2r - Er ** 2, U4))
        with open(self.path("out"), "a") as f:
            f.write("%.3f %.3f %.3f %.3f %.3f %.3f %.3f %.3f %.3f %.3f %.6f\n" % (
                T, h, si, sj, stot, sij, E, E2, Er, E2r, U4))
        for k in range(spinFrame):
            with open(self.path("IsingSpinDistribution.T%.3f.H%.3f.%d.txt" % (T, h, k)), "w") as f:
                f.write("#x       #y       #z       #spin\n")
                fr = frames[k]
                f.write("".join("%.6f %.6f %.6f %.3f\n" % (x, y, z, fr[i])
                                for i, (x, y, z) in enumerate(Lt.pos)))
        return si, sj, sij, auto, E, E2, U4

    def on(self, Lt, r: PointResult, spinFrame: int):
        d = r.data
        T, h = r.Tc, r.H
        si = np.array(d[0:3]); sj = np.array(d[3:6]); sij, auto = d[6], d[7]
        E, E2, U4 = d[8] * T, d[9] * T ** 2, d[10]
        sir = np.array(d[11:14]); sjr = np.array(d[14:17]); sijr = d[17]
        Er, E2r = d[18] * T, d[19] * T ** 2
        zi, zj, zt, hi, hj, ht, Q = d[20:27]
        frames, groups = d[27], d[28]
        vals = (T, h, zi, zj, zt, hi, hj, ht, np.sqrt(si @ si), np.sqrt(sj @ sj), sij, si @ sj, E, E2, U4, sijr,
                sir @ sjr, Er, E2r, Q)
        if self.verbose:
            print("T=%.3f h=%.3f <Siz>=%.3f <Sjz>=%.3f <Sz>=%.3f <Sih>=%.3f <Sjh>=%.3f <Sh>=%.3f <Si>=%.3f "
                  "<Sj>=%.3f <SiSj>=%.3f <Si><Sj>=%.6f <E>=%.3f <E^2>=%.3f <U4>=%.3f <SiSj>r=%.3f "
                  "<Si>r<Sj>r=%.3f <Er>=%.3f <E^2_r>=%.3f <Q>=%.3f" % vals)
        with open(self.path("out"), "a") as f:
            f.write("T= %.6E h= %.6E <Siz>= %.6E <Sjz>= %.6E <Sz>= %.6E <Sih>= %.6E <Sjh>= %.6E <Sh>= %.6E "
                    "<Si>= %.6E <Sj>= %.6E <SiSj>= %.6E <Si><Sj>= %.6E <E>= %.6E <E^2>= %.6E <U4>= %.6E "
                    "<SiSj>r= %.6E <Si>r<Sj>r= %.6E <Er>= %.6E <E^2_r>= %.6E <Q>= %.6E\n" % vals)
        for k in range(spinFrame):
            with open(self.path("OnSpinDistribution.T%.3f.H%.3f.%d.txt" % (T, h, k)), "w") as f:
                f.write("#x       #y       #z       #spinx  #spiny  #spinz\n")
                fr = frames[k]
                f.write("".join("%.6f %.6f %.6f %.6f %.6f %.6f\n" % (x, y, z, fr[i][0], fr[i][1], fr[i][2])
                                for i, (x, y, z) in enumerate(Lt.pos)))
        ng = len(Lt.groups)
        if ng > 0:
            with open(self.path("spinDotSpin.txt"), "a") as f:
                f.write("%.3f %.3f " % (T, h))
                f.write("".join("%.6f " % v for v in groups[: (ng + 1) * (ng + 1) + ng + 1]))
                f.write("\n")
        return si, sj, sij, auto, E, E2, U4, Q

    def result(self, rows):
        with open(self.path("result.txt"), "w") as f:
            f.write("#Temp          #Field         #<Si>          #<Sj>          #Susc          #Energy(K)     "
                    "#Capacity(K/K) #Topo.Q        #U4            #Auto-corr.    \n")
            for row in rows:
                f.write("%15.6E%15.6E%15.6E%15.6E%15.6E%15.6E%15.6E%15.6E%15.6E%15.6E\n" % row)


def start_simulation(rpath: str, outdir: str = ".", seed: int = DEFAULT_SEED, store: str = "f32",
                     tempering: bool = False, dipole_cutoff: Optional[float] = None, verbose: bool = True) -> list:
    """win.startSimulation(updateGUI=False, rpath) with every (T, H) point in
    one batched engine run.  Writes out, spinDotSpin.txt, result.txt and the
    spin-distribution files into outdir; returns the result.txt rows
    (T, H, <Si>, <Sj>, Susc, E, C, Q, U4, autoCorr)."""
    t0 = time.time()
    w = _Writer(outdir, verbose)
    w.headers()
    p = load_input(rpath)
    if p.modelType not in MODELS:
        raise ValueError("for now only Ising, XY and Heisenberg model is supported")
    nc = MODELS[p.modelType]
    Lt = build_input_lattice(p, dipole_cutoff)
    TList = np.linspace(p.T0, p.T1, p.nT)
    HList = np.linspace(p.H0, p.H1, p.nH)
    points = [(float(T), float(H)) for H in HList for T in TList]   # ID = iH*nT + iT (win.py:75-78)
    res = run_points(Lt, nc, p.algorithm, points, p.nthermal, p.nsweep, p.ninterval, p.spinFrame, seed=seed,
                     store=store, tempering=tempering)
    N = Lt.n
    rows = []
    for r in res:
        T = r.T
        if nc == 1:
            si, sj, sij, auto, E, E2, U4 = w.ising(Lt, r, p.spinFrame)
            rows.append((T, r.H, si, sj, (sij - si * sj) / T, E, (E2 - E * E) / T ** 2 * N, 0.0, U4, auto))
        else:
            si, sj, sij, auto, E, E2, U4, Q = w.on(Lt, r, p.spinFrame)
            rows.append((T, r.H, float(np.sqrt(si @ si)), float(np.sqrt(sj @ sj)), (sij - si @ sj) / T, E,
                         (E2 - E * E) / T ** 2 * N, Q, U4, auto))
    w.result(rows)
    if verbose:
        print("time elapsed %.3f s" % (time.time() - t0))
    return rows


def loadMC(rpath: str, **kw) -> list:
    """mcsolver.loadMC (mcsolver/__init__.py:16-17)."""
    return start_simulation(rpath, **kw)


# ---------------------------------------------------------------------------
# mcMain.MC for scripts that drive one point directly (mcsolver/input.py)
# ---------------------------------------------------------------------------

def _bond_entry(b, ising: bool) -> list:
    """A reference-style Bond object (source, target, overLat, strength) or
    a lattice.Bond -> [source, target, overLat, 9 couplings]."""
    if isinstance(b, lat.Bond):
        j9 = b.j9()
    else:
        st = np.atleast_1d(np.asarray(b.strength, np.float64))
        j9 = np.zeros(9)
        j9[: len(st)] = st
    return [int(b.source), int(b.target), [int(x) for x in b.overLat]] + [float(x) for x in j9]


class MC:
    """mcMain.MC (mcMain.py:15-323): one (T, h) point, run on the engine.

    Same constructor arguments; mainLoopViaCLib / mainLoopViaCLib_On return
    what the reference's do and append the same lines to ./out (and write
    the same spin-distribution and spinDotSpin files)."""

    def __init__(self, ID, LMatrix, pos=(), S=(), D=(), bondList=(), T=1, Lx=1, Ly=1, Lz=1, ki_s=0, ki_t=0,
                 ki_overLat=(0, 0, 0), orbGroupList=(), groupInSC=False, h=0., dipoleAlpha=0, On=1, spinFrame=0,
                 localCircuitList=(), outdir=".", seed=DEFAULT_SEED, dipole_cutoff=None):
        norb = len(pos)
        self.ID = ID
        self.Lx, self.Ly, self.Lz, self.norbs = Lx, Ly, Lz, norb
        self.totOrbs = Lx * Ly * Lz * norb
        self.T = lat.clamp_T(T)
        self.h = h
        self.spinFrame = spinFrame
        self.seed = seed
        self._w = _Writer(outdir, True)
        ising = On == 1
        self._p = McInput(LMatrix=[list(map(float, r)) for r in LMatrix], LPack=[Lx, Ly, Lz],
                          pos=[list(map(float, r)) for r in pos], S=list(map(float, S)),
                          DList=[list(map(float, d)) for d in (D if len(D) else [[0, 0, 0]] * norb)],
                          bondList=[_bond_entry(b, ising) for b in bondList], T0=T, T1=T, nT=1, H0=h, H1=h, nH=1,
                          dipoleAlpha=float(dipoleAlpha), GcOrb=[[ki_s, ki_t], list(ki_overLat)],
                          orbGroupList=[list(g) for g in orbGroupList], groupInSC=groupInSC,
                          localCircuitList=list(localCircuitList), spinFrame=spinFrame, nthermal=0, nsweep=0,
                          ninterval=0, xAxisType="T", modelType={1: "Ising", 2: "XY", 3: "Heisenberg"}[On],
                          algorithm="Metropolis")
        self._T_req = T
        self._dc = dipole_cutoff
        self.Lt = build_input_lattice(self._p, dipole_cutoff)

    def _lattice_for(self, On):
        if On == MODELS[self._p.modelType]:
            return self.Lt
        self._p.modelType = {1: "Ising", 2: "XY", 3: "Heisenberg"}[On]
        self.Lt = build_input_lattice(self._p, self._dc)
        return self.Lt

    def mainLoopViaCLib(self, nsweep=1000, nthermal=500, ninterval=-1, algo="Wolff"):
        Lt = self._lattice_for(1)
        r = run_points(Lt, 1, "Wolff" if algo == "Wolff" else "Metropolis", [(self._T_req, self.h)], nthermal,
                       nsweep, ninterval, self.spinFrame, seed=self.seed)[0]
        return self._w.ising(Lt, r, self.spinFrame)

    def mainLoopViaCLib_On(self, nsweep=1000, nthermal=5000, ninterval=-1, algo="Metroplis", On=3, flunc=0.0,
                           h=0., binGraph=False):
        Lt = self._lattice_for(On)
        r = run_points(Lt, On, "Wolff" if algo == "Wolff" else "Metropolis", [(self._T_req, self.h)], nthermal,
                       nsweep, ninterval, self.spinFrame, seed=self.seed, flunc=flunc)[0]
        return self._w.on(Lt, r, self.spinFrame)
