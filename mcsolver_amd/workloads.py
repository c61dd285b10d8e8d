"""The benchmark / parity workloads, built on the engine.

ising_square32          Ising ferromagnet, square L=32 periodic, J=1 (reference input Jxx=-1),
    h=0, 16 temperatures linear in [1.5, 3.5], all-up start, int8 spins.
heis_honeycomb64        Heisenberg honeycomb (CrI3-like, 2-site basis) 64x64 cells: 3 first- and
    6 second-neighbour bonds, each a full 3x3 tensor (diagonal -1 +- 0.1,
    off-diagonal U[-0.05, 0.05], seeded), single-ion Dz = -0.1, 64
    temperatures in [0.1, 2.0]; FP32 or FP16 spin storage.
xy_ladder_scan          XY square lattices L in {16,24,32,48,64,96,128} in one launch, each a
    64-rung parallel-tempering ladder over T in [0.7, 1.2] (swap every
    sweep), Jxx = Jyy = -1 (ferro), FP32 spins, FP64 accumulators.
heis_triangular_dmi128  Heisenberg triangular 128x128, J = -1, Neel DMI D = 0.3 (D = -D0 z x r),
    Dz = -0.05, Bz in [0, 1] x 32 and T in [0.05, 1.0] x 32 (1024 replicas),
    topological charge from 2 triangles per cell.
heis_cubic64            Heisenberg cubic 64^3, J = -1, random start, 128 temperatures in
    [1.2, 1.7]; batched Metropolis, or Wolff on 16 replicas.

Couplings follow the reference's sign convention (negative J is
ferromagnetic, negative D is easy-axis; guiMain.py:42).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List

import numpy as np

from . import _lib as L
from . import lattice as lat
from .engine import Batch, ReplicaSpec, compile_graph

SEED = 0x5EED2026
XY_SCAN_SIZES = (16, 24, 32, 48, 64, 96, 128)


@dataclass
class Workload:
    name: str
    nc: int
    batch: Batch
    lattices: list
    temps: np.ndarray          # per row
    fields: np.ndarray         # per row
    row_graph: np.ndarray      # graph of each row
    store: str
    pt_every: int
    description: str

    @property
    def updates_per_sweep(self) -> int:
        return int(sum(self.batch.graphs[r.graph].n for r in self.batch.reps))


def _graph(nc, Lt, meas_pairs=True):
    J = Lt.J if nc > 1 else Lt.J[:, :, :1]
    return compile_graph(nc, Lt.nbr, J, Lt.S, Lt.D, pairs=Lt.corr_pairs if meas_pairs else None,
                         circuits=Lt.circuits if nc == 3 and len(Lt.circuits) else None)


def honeycomb_full_tensor(L: int, seed: int = SEED) -> lat.Lattice:
    rng = np.random.default_rng(seed)
    bonds = []
    for (s, t, o) in [(0, 1, (0, 0, 0)), (0, 1, (-1, 0, 0)), (0, 1, (0, 1, 0)), (0, 0, (1, 0, 0)), (0, 0, (1, 1, 0)),
                      (0, 0, (0, 1, 0)), (1, 1, (1, 0, 0)), (1, 1, (1, 1, 0)), (1, 1, (0, 1, 0))]:
        J = np.zeros(9)
        J[:3] = -1 + rng.uniform(-0.1, 0.1, 3)
        J[3:] = rng.uniform(-0.05, 0.05, 6)
        bonds.append(lat.Bond(s, t, o, J))
    return lat.build_lattice(L, L, 1, [[1, 0, 0], [-0.5, np.sqrt(3) / 2, 0], [0, 0, 1]],
                             [[1 / 3, 2 / 3, 0], [2 / 3, 1 / 3, 0]], bonds, [1.0, 1.0], [[0, 0, -0.1]] * 2)


def triangular_neel_dmi(L: int, D0: float = 0.3, Dz: float = -0.05) -> lat.Lattice:
    a1, a2 = np.array([1.0, 0, 0]), np.array([0.5, np.sqrt(3) / 2, 0])
    bonds = []
    for o in [(1, 0, 0), (0, 1, 0), (-1, 1, 0)]:
        r = o[0] * a1 + o[1] * a2
        r = r / np.linalg.norm(r)
        Dv = -D0 * np.cross([0, 0, 1.0], r)
        A = np.array([[0, Dv[2], -Dv[1]], [-Dv[2], 0, Dv[0]], [Dv[1], -Dv[0], 0]])
        bonds.append(lat.Bond(0, 0, o, lat.mat_to_tensor9(-np.eye(3) + A)))
    return lat.build_lattice(L, L, 1, [a1, a2, [0, 0, 1]], [[0, 0, 0]], bonds, [1.0], [[0, 0, Dz]],
                             localCircuitList=[((0, (0, 0, 0)), (0, (1, 0, 0)), (0, (0, 1, 0))),
                                               ((0, (1, 0, 0)), (0, (1, 1, 0)), (0, (0, 1, 0)))])


def _init(nc, n, S, mode, rng):
    if nc == 1:
        return np.ones(n)
    if mode == "x":
        v = np.zeros((n, nc))
        v[:, 0] = np.abs(S)
        return v
    v = rng.normal(size=(n, nc))
    return v / np.linalg.norm(v, axis=1, keepdims=True) * np.abs(S)[:, None]


def build(name: str, store: str = "f32", nbins: int = 32, nframes: int = 1, meas: int = 0,
          sizes=XY_SCAN_SIZES, pt: bool = True, nrep_cubic: int = 128, wolff: bool = False, device="cuda",
          seed: int = SEED, recompute_every: int = L.RECOMPUTE_EVERY) -> Workload:
    """The named workload; seed keys its Philox streams (and random starts);
    recompute_every is the period of the full FP64 recompute (drift and
    spin-health checks, 0 = off)."""
    rng = np.random.default_rng(seed)
    reps: List[ReplicaSpec] = []
    graphs, lattices, temps, fields, rgraph, inits = [], [], [], [], [], []
    nc = {"ising_square32": 1, "heis_honeycomb64": 3, "xy_ladder_scan": 2, "heis_triangular_dmi128": 3, "heis_cubic64": 3}[name]
    pt_every = 0
    if name == "ising_square32":
        Lt = lat.square(32, J=-1.0, ising=True)
        graphs.append(_graph(1, Lt)); lattices.append(Lt)
        for k, T in enumerate(np.linspace(1.5, 3.5, 16)):
            reps.append(ReplicaSpec(0, k)); temps.append(T); fields.append(0.0); rgraph.append(0)
            inits.append(_init(1, Lt.n, Lt.S, "x", rng))
        desc = "Ising square L=32, J=1 ferro, h=0, 16 T in [1.5,3.5], int8"
    elif name == "heis_honeycomb64":
        Lt = honeycomb_full_tensor(64)
        graphs.append(_graph(3, Lt)); lattices.append(Lt)
        for k, T in enumerate(np.linspace(0.1, 2.0, 64)):
            reps.append(ReplicaSpec(0, k)); temps.append(T); fields.append(0.0); rgraph.append(0)
            inits.append(_init(3, Lt.n, Lt.S, "x", rng))
        desc = "Heisenberg honeycomb 64x64x2, 3 NN + 6 NNN full-tensor bonds, Dz=-0.1, 64 T in [0.1,2.0], %s" % store
    elif name == "xy_ladder_scan":
        row = 0
        for gi, Ls in enumerate(sizes):
            Lt = lat.square(Ls, J=-1.0)
            graphs.append(_graph(2, Lt)); lattices.append(Lt)
            Ts = np.linspace(0.7, 1.2, 64)
            for k, T in enumerate(Ts):
                if pt:
                    reps.append(ReplicaSpec(gi, row, nrung=64, ladder=gi, rung=k))
                else:
                    reps.append(ReplicaSpec(gi, row + k))
                temps.append(T); fields.append(0.0); rgraph.append(gi)
                inits.append(_init(2, Lt.n, Lt.S, "x", rng))
            row += 64
        pt_every = 1 if pt else 0
        desc = "XY square L in %s mixed in one launch, 64-rung PT ladders T in [0.7,1.2], swap every sweep, fp32" % (
            list(sizes),)
    elif name == "heis_triangular_dmi128":
        Lt = triangular_neel_dmi(128)
        graphs.append(_graph(3, Lt)); lattices.append(Lt)
        k = 0
        for B in np.linspace(0.0, 1.0, 32):
            for T in np.linspace(0.05, 1.0, 32):
                reps.append(ReplicaSpec(0, k)); temps.append(T); fields.append(B); rgraph.append(0)
                inits.append(_init(3, Lt.n, Lt.S, "x", rng)); k += 1
        desc = "Heisenberg triangular 128x128, J=-1, Neel DMI 0.3, Dz=-0.05, 32 Bz x 32 T, %s" % store
    elif name == "heis_cubic64":
        Lt = lat.cubic(64, J=-1.0)
        graphs.append(_graph(3, Lt)); lattices.append(Lt)
        for k, T in enumerate(np.linspace(1.2, 1.7, nrep_cubic)):
            reps.append(ReplicaSpec(0, k)); temps.append(T); fields.append(0.0); rgraph.append(0)
            inits.append(_init(3, Lt.n, Lt.S, "random", rng))
        desc = "Heisenberg cubic 64^3, J=-1, random start, %d T in [1.2,1.7]" % nrep_cubic
    else:
        raise ValueError(name)
    temps = np.asarray(temps)
    beta = 1.0 / np.maximum(temps, 0.1)  # the reference runs T < 0.1 at 0.1 (mcMain.py:21)
    b = Batch(nc, graphs, reps, beta, fields, inits, store=("i8" if nc == 1 else store), meas=meas, nbins=nbins,
              nframes=nframes, seed=seed, wolff=wolff, device=device, recompute_every=recompute_every)
    return Workload(name, nc, b, lattices, temps, np.asarray(fields), np.asarray(rgraph),
                    "i8" if nc == 1 else store, pt_every, desc)


def cubic_replicas(T: float, nrep: int = 16, wolff: bool = False, nframes: int = 1, store: str = "f32",
                   device="cuda") -> Workload:
    """nrep independent replicas of the Heisenberg cubic 64^3 ferromagnet
    (J = -1, random start) at one temperature, each on its own row with one
    bin per frame, so that the bins hold every replica's per-frame series of
    e, e^2, |M|, M^2, M^4 (for autocorrelation times)."""
    rng = np.random.default_rng(SEED + 1)
    Lt = lat.cubic(64, J=-1.0)
    reps = [ReplicaSpec(0, k) for k in range(nrep)]
    inits = [_init(3, Lt.n, Lt.S, "random", rng) for _ in range(nrep)]
    temps = np.full(nrep, float(T))
    b = Batch(3, [_graph(3, Lt)], reps, 1.0 / np.maximum(temps, 0.1), [0.0] * nrep, inits, store=store,
              nbins=nframes, nframes=nframes, seed=SEED, wolff=wolff, device=device,
              recompute_every=L.RECOMPUTE_EVERY)
    return Workload("heis_cubic64", 3, b, [Lt], temps, np.zeros(nrep), np.zeros(nrep, int), store, 0,
                    "Heisenberg cubic 64^3, J=-1, random start, %d replicas at T=%g" % (nrep, T))


# algorithmic HBM bytes per site update: each site's state is read once and
# written once per sweep (couplings are per-lattice tables, Philox is
# counter-based, measurement is fused)
def bytes_per_update(nc: int, store: str) -> int:
    sz = {"f32": 4, "f16": 2, "i8": 1}[store]
    return 2 * nc * sz
