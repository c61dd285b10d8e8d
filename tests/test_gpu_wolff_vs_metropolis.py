"""Wolff cluster updates against batched Metropolis on the cubic Heisenberg
ferromagnet at T = 1.44 (T_c ~ 1.443), the comparison bench.py's
cubic_wolff_vs_metropolis workload makes at 64^3, here at 16^3: per-frame
series of |M| and E of 8 independent replicas per method (one bin per
frame), tau_int from analysis.tau_int, in sweep-equivalents (Metropolis: one
sweep per frame; Wolff: spins flipped per frame / N from the replica state's
flipped-spin count).  Near T_c the cluster update decorrelates the order
parameter in far fewer sweep-equivalents (heisenbergLib.c:310-439 against
:441-473), and both methods measure the same <|M|>.
"""
import numpy as np
import pytest

from mcsolver_amd import _lib as L
from mcsolver_amd import lattice as lat
from mcsolver_amd.analysis import tau_int
from mcsolver_amd.engine import Batch, ReplicaSpec, compile_graph

pytestmark = [pytest.mark.gpu, pytest.mark.slow]
T, NREP, LS = 1.44, 8, 16


def _batch(nframes, wolff, seed):
    Lt = lat.cubic(LS, J=-1.0)
    g = compile_graph(3, Lt.nbr, Lt.J, Lt.S, Lt.D)
    rng = np.random.default_rng(seed)
    spins = []
    for _ in range(NREP):
        v = rng.normal(size=(Lt.n, 3))
        spins.append(v / np.linalg.norm(v, axis=1, keepdims=True))
    b = Batch(3, [g], [ReplicaSpec(0, k) for k in range(NREP)], [1.0 / T] * NREP, [0.0] * NREP, spins,
              nbins=nframes, nframes=nframes, seed=seed, wolff=wolff, recompute_every=L.RECOMPUTE_EVERY)
    b.init()
    return b, Lt.n


def _taus(bins, q):
    return np.array([tau_int(bins[r, :, q])[0] for r in range(NREP)])


def test_wolff_decorrelates_the_order_parameter_faster(gpu):
    # Metropolis: 3000 thermal sweeps, 12000 frames of one sweep
    nm = 12000
    b, n = _batch(nm, False, 11)
    b.run(3000, nm, 1, sweep0=0)
    assert b.error_flag() == 0
    bm = b.bin_data()
    tm = _taus(bm, 2)                       # |M|
    # Wolff from a Metropolis-equilibrated start: 3000 sweeps, 100 cluster
    # steps, then 1500 frames of 8 clusters
    nw, ncl = 1500, 8
    w, _ = _batch(nw, True, 12)
    w.run(3000, 0, 1, sweep0=0)
    w.run(100, 0, 1, algorithm=L.WOLFF, clusters_per_sweep=ncl, sweep0=3000)
    f0 = w.state_data()[:, L.S_ACC].copy()
    w.run(0, nw, 1, algorithm=L.WOLFF, clusters_per_sweep=ncl, sweep0=3100)
    assert w.error_flag() == 0
    flips = w.state_data()[:, L.S_ACC] - f0
    sweq = flips / (n * nw)                 # sweep-equivalents per frame, per replica
    bw = w.bin_data()
    tw = _taus(bw, 2) * sweq
    mm, mw = bm[:, :, 2].mean(axis=1), bw[:, :, 2].mean(axis=1)
    print("tau_int(|M|): Metropolis %.1f +- %.1f sweeps, Wolff %.2f +- %.2f sweep-eq (%.2f sweep-eq per frame)" %
          (tm.mean(), tm.std(ddof=1) / np.sqrt(NREP), tw.mean(), tw.std(ddof=1) / np.sqrt(NREP), sweq.mean()))
    print("<|M|>: Metropolis %.4f +- %.4f, Wolff %.4f +- %.4f" %
          (mm.mean(), mm.std(ddof=1) / np.sqrt(NREP), mw.mean(), mw.std(ddof=1) / np.sqrt(NREP)))
    assert 0.05 < sweq.mean() < 50
    assert tw.mean() * 4 < tm.mean()
    se = np.hypot(mm.std(ddof=1), mw.std(ddof=1)) / np.sqrt(NREP)
    assert abs(mm.mean() - mw.mean()) < 4.5 * se + 2e-3
