"""The benchmark configurations at their own lattices, against exact results
and the FP64 oracle (tests/golden/config_fp64.json, written by
tests/golden/make_config_fp64.py: the oracle restatement of the reference's
MCMainFunction, random-sequential Metropolis with glibc rand(), R seeds per
point).

Each comparison is between independent runs: R replicas per point on the
GPU (distinct Philox keys) and R oracle seeds.  welch_z gives the normal
deviate of the difference of the run means; every |z| must stay below 4.5
and the chi-square of a test (sum z^2 over its comparisons) below its 99.9 %
quantile.  Observables (reference definitions: win.py:97-101,136-142,
mcMain.py:251-253):
  E per site, C = N (<e^2> - <e>^2) / T^2, |M| = |total spin| / N,
  U4 = <M^2>^2 / <M^4>, Q = sum of triangle solid angles / 4 pi,
  Ising chi = (spin_ij - spin_i spin_j) / T from the correlation pairs.
The GPU runs use unscaled couplings and beta = 1/T (energies in energy
units); the oracle's tuple is in beta units (E x T, E2 x T^2).
"""
import json
import os

import numpy as np
import pytest

import exact
from stats import chi2_summary, welch_z
from mcsolver_amd import _lib as L
from mcsolver_amd import lattice as lat
from mcsolver_amd import workloads as W
from mcsolver_amd.dropin import ising_observables
from mcsolver_amd.engine import Batch, ReplicaSpec, compile_graph

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "config_fp64.json")
ZMAX = 4.5


def _fix(section):
    return json.load(open(FIX))[section]


def _check(rows, label):
    """rows: (name, z, text).  Prints the table and the chi-square summary."""
    for _, z, txt in rows:
        print("%s  z=%+.2f" % (txt, z))
    zs = [z for _, z, _ in rows]
    chi2, dof, lim, f2 = chi2_summary(zs)
    print("%s: chi2 = %.1f for %d comparisons (99.9%% quantile %.1f); %.0f%% within 2 sigma" %
          (label, chi2, dof, lim, 100 * f2), flush=True)
    bad = [txt for _, z, txt in rows if not abs(z) < ZMAX]
    assert not bad, "beyond %.1f sigma:\n" % ZMAX + "\n".join(bad)
    assert chi2 < lim, "chi2 %.1f >= %.1f" % (chi2, lim)


def _batch(nc, Lt, points, R, nframes, store="f32", meas=0, init="x", seed=0x5EED2026, wolff=False, **kw):
    """R independent replicas at each (T, h) point, one row each."""
    J = Lt.J if nc > 1 else Lt.J[:, :, :1]
    g = compile_graph(nc, Lt.nbr, J, Lt.S, Lt.D, pairs=Lt.corr_pairs,
                      circuits=Lt.circuits if nc == 3 and len(Lt.circuits) else None)
    rng = np.random.default_rng(seed)
    reps, beta, hs, inits = [], [], [], []
    for T, h in points:
        for _ in range(R):
            reps.append(ReplicaSpec(0, len(reps)))
            beta.append(1.0 / max(T, 0.1))
            hs.append(h)
            if nc == 1:
                inits.append(np.ones(Lt.n))
            elif init == "x":
                v = np.zeros((Lt.n, nc))
                v[:, 0] = np.abs(Lt.S)
                inits.append(v)
            else:
                v = rng.normal(size=(Lt.n, nc))
                inits.append(v / np.linalg.norm(v, axis=1, keepdims=True))
    b = Batch(nc, [g], reps, beta, hs, inits, store=store, meas=meas, nbins=1, nframes=nframes, seed=seed,
              wolff=wolff, **kw)
    b.init()
    return b


def _per_run(b, npts, R):
    """[npts, R] arrays of per-run means from the row accumulators."""
    A = b.accumulators()
    n = A[:, L.A_COUNT]
    out = {k: (A[:, i] / n).reshape(npts, R) for k, i in
           (("E", L.A_E), ("E2", L.A_E2), ("M", L.A_M), ("M2", L.A_M2), ("M4", L.A_M4), ("Q", L.A_Q))}
    return out, A


def _on_rows(label, pts, g, fx, N, names, rel_floor=0.0, q_floor=1e-6):
    """Compare E, C, |M|, U4 (and Q) of GPU runs g[point] with oracle runs."""
    rows = []
    for k, (key, p) in enumerate(pts):
        T = p["T"]
        o = np.asarray(p["runs"])                       # E, E2, absM, Q, M2, M4 (beta units)
        oE, oE2 = o[:, 0] * T, o[:, 1] * T * T
        ovals = {"E": oE, "C": N * (oE2 - oE * oE) / T ** 2, "M": o[:, 2], "U4": o[:, 4] ** 2 / o[:, 5],
                 "Q": o[:, 3]}
        gE, gE2 = g["E"][k], g["E2"][k]
        gvals = {"E": gE, "C": N * (gE2 - gE * gE) / T ** 2, "M": g["M"][k], "U4": g["M2"][k] ** 2 / g["M4"][k],
                 "Q": g["Q"][k]}
        for nm in names:
            z = welch_z(gvals[nm], ovals[nm], rel_floor, q_floor if nm == "Q" else 0.0)
            rows.append((nm, z, "%s %-10s T=%.4f h=%.3f %-3s gpu %.6g  oracle %.6g" %
                         (label, key, T, p["h"], nm, np.mean(gvals[nm]), np.mean(ovals[nm]))))
    return rows


# ---------------------------------------------------------------------------
# Ising 32x32, 16 temperatures, 2e4 + 5e4 sweeps
# ---------------------------------------------------------------------------
def test_ising_square32_vs_oracle(gpu):
    fx = _fix("ising_square32")
    nth, nsw = fx["nthermal"], fx["nsweep"]
    Lt = lat.square(32, J=-1.0, ising=True)
    temps = np.linspace(1.5, 3.5, 16)
    R = 8
    b = _batch(1, Lt, [(T, 0.0) for T in temps], R, nsw, meas=L.MEAS_PAIRS)
    b.run(nth, nsw, 1, sweep0=0)
    assert b.error_flag() == 0
    g, A = _per_run(b, len(temps), R)
    N = Lt.n
    rows = []
    # E and C against Kaufman's exact finite-torus values, every temperature
    for k, T in enumerate(temps):
        e_ex, c_ex = exact.ising_e_c(32, T)
        c = N * (g["E2"][k] - g["E"][k] ** 2) / T ** 2
        for nm, v, ex in (("E", g["E"][k], e_ex), ("C", c, c_ex)):
            se = v.std(ddof=1) / np.sqrt(R)
            z = welch_z(v, [ex, ex], 0.0, 1e-12)  # the exact value has no spread
            rows.append((nm, z, "ising32 T=%.4f %-2s gpu %.6f +- %.2g  exact %.6f" % (T, nm, v.mean(), se, ex)))
    # U4 and chi against the FP64 oracle
    S = np.ones(N)
    for key, p in fx["points"].items():
        T = p["T"]
        k = int(np.argmin(np.abs(temps - T)))
        assert abs(temps[k] - T) < 1e-9
        o = np.asarray(p["runs"])
        o_u4, o_chi = o[:, 8], (o[:, 2] - o[:, 0] * o[:, 1]) / T
        gv = np.array([ising_observables(A[k * R + r], nsw, S, N, False, N) for r in range(R)])
        g_u4, g_chi = gv[:, 8], (gv[:, 2] - gv[:, 0] * gv[:, 1]) / T
        for nm, gg, oo in (("U4", g_u4, o_u4), ("chi", g_chi, o_chi)):
            z = welch_z(gg, oo)
            rows.append((nm, z, "ising32 T=%.4f %-3s gpu %.6f  oracle %.6f" % (T, nm, gg.mean(), oo.mean())))
    _check(rows, "Ising 32^2 x 16 T")


# ---------------------------------------------------------------------------
# triangular 128x128 Neel DMI, on the FP32 streaming path and
# the FP16 LDS path
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("store", ["f32", "f16"])
def test_triangular_dmi128_vs_oracle(gpu, store):
    fx = _fix("tri_dmi128")
    nth, nsw = fx["nthermal"], fx["nsweep"]   # the oracle's schedule: same run-length bias of C and U4
    Lt = W.triangular_neel_dmi(128)
    pts = list(fx["points"].items())
    R = fx["runs_per_point"]
    b = _batch(3, Lt, [(p["T"], p["h"]) for _, p in pts], R, nsw, store=store, meas=L.MEAS_PAIRS | L.MEAS_Q)
    b.run(nth, nsw, 1, sweep0=0)
    assert b.error_flag() == 0
    g, _ = _per_run(b, len(pts), R)
    rows = _on_rows("tri128 %s" % store, pts, g, fx, Lt.n, ("E", "C", "M", "U4", "Q"),
                    rel_floor=1e-4 if store == "f32" else 4e-3)
    _check(rows, "triangular DMI 128^2 (%s)" % store)


# ---------------------------------------------------------------------------
# cubic Heisenberg, Metropolis and Wolff at 16^3 around T_c;
# the 64^3 lattice for spin lengths and the FP64 energy check
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("algorithm", ["metropolis", "wolff"])
def test_cubic16_vs_oracle(gpu, algorithm):
    fx = _fix("cubic16")
    Lt = lat.cubic(16, J=-1.0)
    pts = list(fx["points"].items())
    R = 8
    if algorithm == "metropolis":
        nth, nsw = fx["nthermal"], fx["nsweep"]
        b = _batch(3, Lt, [(p["T"], 0.0) for _, p in pts], R, nsw, init="random")
        b.run(nth, nsw, 1, sweep0=0)
    else:
        nth, nsw, ncl = 300, 3000, 8
        b = _batch(3, Lt, [(p["T"], 0.0) for _, p in pts], R, nsw, init="random", wolff=True)
        b.run(nth, nsw, 1, algorithm=L.WOLFF, clusters_per_sweep=ncl, sweep0=0)
    assert b.error_flag() == 0
    g, _ = _per_run(b, len(pts), R)
    rows = _on_rows("cubic16 %s" % algorithm, pts, g, fx, Lt.n, ("E", "C", "M", "U4"))
    _check(rows, "cubic 16^3 %s" % algorithm)


def test_cubic64_lengths_and_energy_check(gpu):
    """64^3 (streaming path, 3 MB per replica): after 200 Metropolis sweeps
    with the FP64 recompute every 25 sweeps, every spin has unit length
    (one Newton step per accepted update) and the tracked energy never
    drifted from the recomputed one by more than 1e-7 of |E|; a few Wolff
    clusters keep unit lengths too."""
    Lt = lat.cubic(64, J=-1.0)
    temps = [1.2, 1.44, 1.7]
    b = _batch(3, Lt, [(T, 0.0) for T in temps], 1, 1, init="random", recompute_every=25)
    b.run(200, 0, 1, sweep0=0)
    st = b.state_data()
    assert b.error_flag() == 0
    for k in range(len(temps)):
        s = b.replica_spins(k)
        dev = np.abs(np.linalg.norm(s.astype(np.float64), axis=1) - 1.0).max()
        print("T=%.2f E=%.4f drift=%.3g max||s|-1|=%.2g" % (temps[k], st[k, L.S_E], st[k, L.S_DRIFT], dev))
        assert dev < 2e-6
        assert st[k, L.S_DRIFT] < 1e-7 * abs(st[k, L.S_E])
        assert st[k, L.S_DRIFT] > 0.0  # the check ran
    w = _batch(3, Lt, [(1.44, 0.0)], 2, 1, init="random", wolff=True)
    w.run(4, 0, 1, algorithm=L.WOLFF, clusters_per_sweep=8, sweep0=0)
    assert w.error_flag() == 0
    for k in range(2):
        s = w.replica_spins(k)
        assert np.abs(np.linalg.norm(s.astype(np.float64), axis=1) - 1.0).max() < 2e-6
