"""Batched XY ladder scan (square lattices, 64-rung PT ladders) vs the FP64 oracle (reference algorithm).

The batched persistent engine (LDS kernel, PT every sweep) and the plain
one-launch-per-replica-per-colour path run the same trajectory bit for bit
(test_gpu_bitexact.py), so one comparison covers both: U4(T) and C(T) of
the batched run against the random-sequential FP64 oracle restatement of
xyLib.c at the same L and T (8 independent oracle runs per point).
Criterion: at least 75 % of the compared points within 2 combined sigma and
none beyond 5 (the oracle's sigma comes from 8 runs, so its tails are t-like).
"""
import multiprocessing as mp
import os

import numpy as np
import pytest

from mcsolver_amd import _lib as L
from mcsolver_amd import lattice as lat
from mcsolver_amd import workloads as W
from mcsolver_amd.analysis import batch_estimates
from oracle import mcref

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

SIZES = (16, 32)
NTH, NSW = 4000, 40000
T_CHECK = [0, 13, 26, 38, 51, 63]   # rungs compared against the oracle


def _oracle_point(args):
    L_, T, nsw, seed = args
    Lt = lat.square(L_, J=-1.0)
    a = lat.on_args(Lt, T, 0.0, 2000, nsw)
    out, _, _ = mcref.on_run(2, a, seed=seed)
    e, e2, u4 = out[8] * T, out[9] * T * T, out[10]
    return L_, T, u4, Lt.n * (e2 - e * e) / T ** 2


def test_xy_scan_batched_matches_oracle(gpu):
    wl = W.build("xy_ladder_scan", sizes=SIZES, nbins=20, nframes=NSW)
    b = wl.batch
    b.init()
    b.run(NTH, NSW, 1, pt_every=1)
    assert b.error_flag() == 0
    n_sites = [wl.lattices[g].n for g in wl.row_graph]
    est = batch_estimates(b, wl.temps, n_sites)
    pts = []
    for gi, L_ in enumerate(SIZES):
        for t in T_CHECK:
            row = gi * 64 + t
            for s in range(8):
                pts.append((L_, float(wl.temps[row]), 12000, 100 * row + s))
    P = min(16, len(os.sched_getaffinity(0)))
    with mp.get_context("fork").Pool(P) as pool:
        res = pool.map(_oracle_point, pts)
    within2, n = 0, 0
    for gi, L_ in enumerate(SIZES):
        for t in T_CHECK:
            row = gi * 64 + t
            T = float(wl.temps[row])
            rs = np.array([(r[2], r[3]) for r in res if r[0] == L_ and abs(r[1] - T) < 1e-12])
            for k, (name, err) in enumerate((("U4", "U4_err"), ("C", "C_err"))):
                om, oe = rs[:, k].mean(), rs[:, k].std(ddof=1) / np.sqrt(len(rs))
                g, ge = est[row][name], est[row][err]
                z = abs(g - om) / np.hypot(ge, oe)
                print("L=%3d T=%.4f %-2s gpu %.5f+-%.5f  oracle %.5f+-%.5f  z=%.2f" % (L_, T, name, g, ge, om, oe, z))
                assert z < 5.0, (L_, T, name)
                within2 += z < 2.0
                n += 1
    print("within 2 sigma: %d / %d" % (within2, n))
    assert within2 >= 0.75 * n
    # tempering moved replicas along every ladder
    acc = b.pt_acc.cpu().numpy()
    for gi in range(len(SIZES)):
        assert (acc[gi * 64:gi * 64 + 63] > 0).all()
