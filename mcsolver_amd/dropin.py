"""Drop-in implementation of the reference's MCMainFunction entry points.

``isinglib.MCMainFunction`` (isingLib.c:259-451) and
``{xylib,heisenberglib}.MCMainFunction`` (xyLib.c:413-809,
heisenbergLib.c:478-885) take the flattened per-site tuples that
mcMain.py:119-124 / 239-248 build and return an 11-tuple / 29-tuple.  Here
the tuples are compiled into device tables (engine.compile_graph) and one
replica runs on the MI355X through libmcgpu; the returned tuples have the
reference's shapes, order, sentinels and observable definitions.

Schedule mapping.  The reference does `ninterval` random single-site updates
between measurements (ninterval = N by default, mcMain.py:54, 145); the
engine does sweeps_per_frame = max(1, round(ninterval/N)) colour-ordered
sweeps, and nthermal*sweeps_per_frame thermalisation sweeps.  Wolff does
`ninterval` clusters per measurement as the reference does.

Reproduced observable conventions (they define the reference's numbers):
Ising couplings read from linkStrength[i + l] (isingLib.c:30); Ising
Metropolis energy relative to the end of thermalisation (isingLib.c:359);
Ising spin_tot started from the first `ninterval` spins (isingLib.c:350) and,
under Wolff, blind to flips of the last site (isingLib.c:228); U4 =
<M^2>^2/<M^4>; XY field and 'h' projections on x (xyLib.c:203,633); XY
M = |S_i|/nLat (xyLib.c:654); Heisenberg M = |totSpin|/nLat
(heisenbergLib.c:726).  Errors raise instead of crashing (the reference
ignores PyArg_ParseTuple's status, isingLib.c:277).
"""
from __future__ import annotations

from typing import Optional

import numpy as np

from . import _lib as L
from .engine import Batch, ReplicaSpec, compile_graph, reciprocal_couplings

DEFAULT_SEED = 0x5EED2026


def _plan_counts(n: int, nthermal: int, nsweep: int, ninterval: int, wolff: bool):
    if ninterval <= 0:
        raise ValueError("ninterval must be positive (mcMain.py passes N when it is <= 0)")
    if wolff:
        return nthermal, 1, ninterval
    spf = max(1, int(round(ninterval / n)))
    return nthermal * spf, spf, 0


def _run_frames(b: Batch, nth: int, nsweep: int, spf: int, alg: int, ncl: int, spinFrame: int,
                replicas=(0,)) -> list:
    """Run thermalisation + nsweep frames; collect spin snapshots of the given
    replicas at frames 0, ops, 2*ops, ... with ops = nsweep // spinFrame
    (isingLib.c:366-386).  Returns frames[k] = list of [n, nc] arrays of
    replica replicas[k].  Philox counters depend only on the sweep index, so
    splitting the run at the snapshots leaves the trajectory unchanged."""
    frames = [[] for _ in replicas]
    if spinFrame <= 0:
        b.run(nth, nsweep, spf, algorithm=alg, clusters_per_sweep=ncl)
        return frames
    ops = nsweep // spinFrame
    if ops <= 0:
        raise ValueError("spinFrame > nsweep (the reference divides by zero here)")
    marks = list(range(0, nsweep, ops))
    done = 0
    sweep0 = 0
    first = True
    for m in marks + [nsweep]:
        todo = m + 1 - done if m < nsweep else nsweep - done
        if todo > 0:
            th = nth if first else 0
            b.run(th, todo, spf, algorithm=alg, clusters_per_sweep=ncl, sweep0=sweep0, frame0=done)
            sweep0 += th + todo * spf
            done += todo
            first = False
        if m < nsweep and len(frames[0]) < spinFrame:
            for k, r in enumerate(replicas):
                frames[k].append(b.replica_spins(r).copy())
    return frames


def recompute_period(nbr, J, ising: bool, diag_only: bool = False) -> int:
    """Period of the full FP64 recompute for a run of these couplings: the
    production period when the couplings are reciprocal, 0 (no recompute)
    otherwise -- then no energy function has the tracked changes, and the
    reference's incrementally tracked energy is the quantity it reports."""
    return L.RECOMPUTE_EVERY if reciprocal_couplings(nbr, J, ising, diag_only) else 0


def check_health(b: Batch) -> None:
    """Raise if the engine's device error flag is set: a persistent launch's
    barrier timed out, or a full FP64 recompute found NaN / Inf in the spins
    or spin lengths off their S (include/mcgpu.h MCG_ERR_*)."""
    f = b.error_flag()
    if f:
        what = [nm for bit, nm in ((L.ERR_BARRIER, "grid barrier timed out"),
                                   (L.ERR_NONFINITE, "non-finite energy or magnetisation (NaN/Inf spins)"),
                                   (L.ERR_LENGTH, "spin lengths off their S"),
                                   (L.ERR_DRIFT, "tracked energy off the full recompute")) if f & bit]
        raise RuntimeError("engine error flag 0x%x: %s" % (f, "; ".join(what) or "unknown"))


def ising_couplings(linkStrength, n: int, maxl: int, nlink, per_link_couplings: bool = False) -> np.ndarray:
    """[n, maxl] coupling of link l of site i: entry i + l of the flat table,
    as isingLib.c:30 reads it, or entry i*maxl + l with per_link_couplings.
    Links beyond nlink[i] get 0."""
    flat = np.asarray(linkStrength, np.float64).reshape(-1)
    nl = np.asarray(nlink, np.int64)
    lmask = np.arange(maxl)[None, :] < nl[:, None]
    if per_link_couplings:
        idx = np.arange(n)[:, None] * maxl + np.arange(maxl)[None, :]
    else:
        idx = np.arange(n)[:, None] + np.arange(maxl)[None, :]
    return np.where(lmask, flat[np.clip(idx, 0, flat.size - 1)], 0.0)


def ising_observables(A: np.ndarray, nsweep: int, S: np.ndarray, ninterval: int, wolff: bool, nLat: int) -> tuple:
    """The first ten entries of isinglib's result (isingLib.c:436-446) from
    one row of engine accumulators: spin_i, spin_j, spin_ij, autoCorr, E, E2,
    E_rnorm, E2_rnorm, U4, spin_tot.  S = initial spins (all +S)."""
    ns = float(nsweep)
    n = len(S)
    # reference totSpin starts at the sum of the first ninterval spins
    # (isingLib.c:350) and, under Wolff, misses flips of the last site
    # (isingLib.c:228)
    off = S[: min(int(ninterval), n)].sum() - S.sum()
    tot = A[L.A_TOT] / ns + off
    if wolff:
        tot += -A[L.A_LAST] / ns + S[n - 1]
    U4 = (A[L.A_M2] / ns) * (A[L.A_M2] / ns) / (A[L.A_M4] / ns)
    return (A[L.A_SI] / ns, A[L.A_SJ] / ns, A[L.A_SIJ] / ns, A[L.A_MM] / ns - (A[L.A_M] / ns) * (A[L.A_M] / ns),
            A[L.A_E] / ns, A[L.A_E2] / ns, A[L.A_ER] / ns, A[L.A_ER2] / ns, U4, tot / nLat)


def on_observables(A: np.ndarray, nsweep: int, nc: int, has_rnorm: bool) -> list:
    """The first 27 entries of {xy,heisenberg}lib's result
    (heisenbergLib.c:854-880, xyLib.c:781-805) from one row of engine
    accumulators."""
    ns = float(nsweep)
    three = (lambda base: [A[base] / ns, A[base + 1] / ns, A[base + 2] / ns if nc == 3 else 0.0])
    Mb = A[L.A_M] / ns
    auto = A[L.A_MM] / ns - Mb * Mb
    U4 = (A[L.A_M2] / ns) * (A[L.A_M2] / ns) / (A[L.A_M4] / ns)
    vals = three(L.A_SI) + three(L.A_SJ) + [A[L.A_SIJ] / ns, auto, A[L.A_E] / ns, A[L.A_E2] / ns, U4] + \
        three(L.A_SIR) + three(L.A_SJR) + [A[L.A_SIJR] / ns, A[L.A_ER] / ns, A[L.A_ER2] / ns] + \
        [A[L.A_ZI] / ns, A[L.A_ZJ] / ns, A[L.A_ZT] / ns, A[L.A_HI] / ns, A[L.A_HJ] / ns, A[L.A_HT] / ns,
         A[L.A_Q] / ns if nc == 3 else 0.0]
    if not has_rnorm:
        # no block-spin sites: the reference divides 0/0 (heisenbergLib.c:784-790)
        for k in range(11, 20):
            vals[k] = float("nan")
    return [float(v) for v in vals]


def on_initial_spins(S: np.ndarray, nc: int, flunc: float = 0.0, seed: int = DEFAULT_SEED) -> np.ndarray:
    """Initial O(n) spins: along x with length S (sign kept), tilted by a
    random vector of length flunc and renormalised (heisenbergLib.c:160-172,
    xyLib.c:120-132).  mcMain always passes flunc = 0."""
    n = len(S)
    init = np.zeros((n, nc))
    init[:, 0] = S
    if float(flunc) != 0.0:
        rng = np.random.default_rng(seed)
        r = rng.normal(size=(n, nc))
        r /= np.linalg.norm(r, axis=1, keepdims=True)
        v = init + float(flunc) * r
        ln = np.linalg.norm(v, axis=1, keepdims=True)
        v = np.where(ln >= 1e-5, v / np.where(ln > 0, ln, 1), v)
        init = v * np.abs(S)[:, None]
    return init


def ising_main(algorithm, initSpin, nthermal, nsweep, ninterval, maxNLinking, nlink, linkStrength, linkedOrb,
               corrOrbPair, h, rOrb, rOrbCluster, linkedOrb_rnorm, spinFrame, callback=None,
               seed: int = DEFAULT_SEED, per_link_couplings: bool = False):
    """isinglib.MCMainFunction on the GPU.

    By default link l of site i takes coupling linkStrength[i + l], as the
    reference reads it (isingLib.c:30).  With equal couplings on every link
    that is the same as linkStrength[i*maxNLinking + l]; with unequal ones
    the reference's update rule is not derived from any Hamiltonian and its
    stationary state depends on the update order, so no coloured schedule
    reproduces it.  per_link_couplings=True uses linkStrength[i*maxNLinking
    + l], the Hamiltonian the input describes.
    """
    S = np.asarray(initSpin, np.float64)
    n = len(S)
    maxl = int(maxNLinking)
    nl = np.asarray(nlink, np.int64)
    nbr = np.asarray(linkedOrb, np.int64).reshape(n, maxl)
    if len(linkStrength) != n * maxl or nl.size != n:
        raise ValueError("inconsistent tuple lengths")
    nbr = np.where(np.arange(maxl)[None, :] < nl[:, None], nbr, -1)
    Jeff = ising_couplings(linkStrength, n, maxl, nl, per_link_couplings)
    pairs = np.asarray(corrOrbPair, np.int64).reshape(-1, 2)
    nr = len(rOrb)
    rclu = np.asarray(rOrbCluster, np.int64).reshape(nr, -1) if nr else None
    rl = np.asarray(linkedOrb_rnorm, np.int64).reshape(nr, maxl) if nr else None
    g = compile_graph(1, nbr, Jeff[:, :, None], S, np.zeros((n, 3)), pairs=pairs, rOrb=rOrb, rOrbCluster=rclu,
                      rlink=rl)
    wolff = int(algorithm) == 1
    alg = L.WOLFF if wolff else L.METROPOLIS
    nth, spf, ncl = _plan_counts(n, int(nthermal), int(nsweep), int(ninterval), wolff)
    meas = ising_meas_flags(nr > 0, wolff)
    # sigma = +1 everywhere: the spin of site i is sigma_i * initSpin[i]
    b = Batch(1, [g], [ReplicaSpec(0, 0)], [1.0], [float(h)], [np.ones(n)], meas=meas, nbins=1,
              nframes=int(nsweep), seed=seed, wolff=wolff, recompute_every=recompute_period(nbr, Jeff, True))
    b.init()
    frames = _run_frames(b, nth, int(nsweep), spf, alg, ncl, int(spinFrame))[0]
    check_health(b)
    out = ising_observables(b.accumulators()[0], int(nsweep), S, int(ninterval), wolff, len(pairs))
    if int(spinFrame) > 0:
        fr = tuple(tuple(float(v) for v in (f[:, 0] * S)) for f in frames)
    else:
        fr = 0.0
    return tuple(float(v) for v in out) + (fr,)


def ising_meas_flags(has_rnorm: bool, wolff: bool) -> int:
    """Measurement flags of the Ising result: pairs, block spins, and the
    Metropolis energy relative to the end of thermalisation (isingLib.c:359)."""
    return L.MEAS_PAIRS | (L.MEAS_RNORM if has_rnorm else 0) | (0 if wolff else L.ISING_RELE)


def on_meas_flags(nc: int, has_circuits: bool, has_rnorm: bool, has_groups: bool) -> int:
    return L.MEAS_PAIRS | (L.MEAS_Q if (nc == 3 and has_circuits) else 0) | \
        (L.MEAS_RNORM if has_rnorm else 0) | (L.MEAS_GROUPS if has_groups else 0)


def on_main(nc: int, algorithm, initSpin, initD, nthermal, nsweep, ninterval, maxNLinking, nlink, linkStrength,
            linkedOrb, localCircuits, corrOrbPair, nOrbGroup, maxOrbGroupSize, orbGroupList, flunc, h, rOrb,
            rOrbCluster, linkedOrb_rnorm, spinFrame, ignoreNonDiagonalJ, callback=None, seed: int = DEFAULT_SEED,
            store: str = "f32"):
    S = np.asarray(initSpin, np.float64)
    n = len(S)
    maxl = int(maxNLinking)
    nl = np.asarray(nlink, np.int64)
    nbr = np.asarray(linkedOrb, np.int64).reshape(n, maxl)
    J = np.asarray(linkStrength, np.float64).reshape(n, maxl, 9)
    D = np.asarray(initD, np.float64).reshape(n, 3)
    lmask = np.arange(maxl)[None, :] < nl[:, None]
    nbr = np.where(lmask, nbr, -1)
    pairs = np.asarray(corrOrbPair, np.int64).reshape(-1, 2)
    circ = np.asarray(localCircuits, np.int64).reshape(-1, 3) if nc == 3 else None
    nr = len(rOrb)
    rclu = np.asarray(rOrbCluster, np.int64).reshape(nr, -1) if nr else None
    rl = np.asarray(linkedOrb_rnorm, np.int64).reshape(nr, maxl) if nr else None
    ng = int(nOrbGroup)
    groups = []
    if ng > 0:
        gl = np.asarray(orbGroupList, np.int64).reshape(ng, int(maxOrbGroupSize))
        for row in gl:
            m = []
            for v in row:
                if v < 0:
                    break
                m.append(int(v))
            groups.append(m)
    g = compile_graph(nc, nbr, J, S, D, pairs=pairs, circuits=circ, rOrb=rOrb, rOrbCluster=rclu, rlink=rl,
                      groups=groups, diag_only=int(ignoreNonDiagonalJ) > 0)
    init = on_initial_spins(S, nc, float(flunc), seed)
    wolff = int(algorithm) == 1
    alg = L.WOLFF if wolff else L.METROPOLIS
    nth, spf, ncl = _plan_counts(n, int(nthermal), int(nsweep), int(ninterval), wolff)
    meas = on_meas_flags(nc, circ is not None and len(circ) > 0, nr > 0, ng > 0)
    b = Batch(nc, [g], [ReplicaSpec(0, 0)], [1.0], [float(h)], [init], store=store, meas=meas, nbins=1,
              nframes=int(nsweep), seed=seed, wolff=wolff,
              recompute_every=recompute_period(nbr, J, False, int(ignoreNonDiagonalJ) > 0))
    b.init()
    frames = _run_frames(b, nth, int(nsweep), spf, alg, ncl, int(spinFrame))[0]
    check_health(b)
    vals = on_observables(b.accumulators()[0], int(nsweep), nc, nr > 0)
    if int(spinFrame) > 0:
        fr = tuple(tuple((float(v[0]), float(v[1]), float(v[2]) if nc == 3 else 0.0) for v in f) for f in frames)
    else:
        fr = 0.0
    if ng > 0:
        gr = tuple(float(v) / float(nsweep) for v in b.group_acc()[0])
    else:
        gr = 0.0
    return tuple(vals) + (fr, gr)
