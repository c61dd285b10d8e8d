"""The batched XY ladder scan (square lattices L = 16 ... 128 in
one launch, 64-rung parallel-tempering ladders over T in [0.7, 1.2], swap
every sweep), against the FP64 oracle.

The GPU side is the benchmark's own workload and launch plan (persistent
launch with tempering inside, the L = 128 replicas split over workgroups),
run R times with independent Philox keys.  The oracle side is
tests/golden/config_fp64.json (sections xy_square_small: L = 16, 32 at 6
rungs; xy_square_large: L = 64 at 3 rungs, L = 128 at 2 rungs), R seeds of
the random-sequential FP64 restatement of xyLib.c each (no tempering: the
equilibrium values are the same).  Every comparison uses independent runs on
both sides; |z| < 4.5 each, and the chi-square over all comparisons below
its 99.9 % quantile.  The plain one-launch-per-replica-per-colour path runs
the same trajectory bit for bit (test_gpu_bitexact.py), so this covers both
sides of the batched / plain speedup.
"""
import json
import os

import numpy as np
import pytest

from stats import chi2_summary, welch_z
from mcsolver_amd import _lib as L
from mcsolver_amd import workloads as W

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "config_fp64.json")
NTH, NSW, R = 4000, 12000, 8


def test_xy_scan_batched_matches_oracle(gpu):
    d = json.load(open(FIX))
    pts = {}
    for sec in ("xy_square_small", "xy_square_large"):
        assert d[sec]["nsweep"] == NSW
        pts.update(d[sec]["points"])
    runs = {}
    for s in range(R):
        wl = W.build("xy_ladder_scan", nbins=1, nframes=NSW, seed=0x5EED2026 + 7919 * s)
        b = wl.batch
        b.init()
        b.run(NTH, NSW, 1, pt_every=1, sweep0=0)
        assert b.error_flag() == 0
        A = b.accumulators()
        acc = b.pt_acc.cpu().numpy()
        for gi in range(len(W.XY_SCAN_SIZES)):
            assert (acc[gi * 64:gi * 64 + 63] > 0).all(), "a ladder pair never swapped"
        for key, p in pts.items():
            L_, r = (int(x[1:]) for x in key.split("_"))
            row = W.XY_SCAN_SIZES.index(L_) * 64 + r
            assert abs(wl.temps[row] - p["T"]) < 1e-9
            a = A[row] / A[row, L.A_COUNT]
            runs.setdefault(key, []).append([a[L.A_E], a[L.A_E2], a[L.A_M], a[L.A_M2], a[L.A_M4]])
        print("run %d done" % s, flush=True)
        b.close()
        del wl, b
    zs = []
    for key, p in pts.items():
        L_ = int(key.split("_")[0][1:])
        N, T = L_ * L_, p["T"]
        o = np.asarray(p["runs"])                    # E, E2, absM, Q, M2, M4 (beta units)
        oE, oE2 = o[:, 0] * T, o[:, 1] * T * T
        g = np.asarray(runs[key])
        gE, gE2 = g[:, 0], g[:, 1]
        for nm, gv, ov in (("E", gE, oE), ("C", N * (gE2 - gE * gE) / T ** 2, N * (oE2 - oE * oE) / T ** 2),
                           ("M", g[:, 2], o[:, 2]), ("U4", g[:, 3] ** 2 / g[:, 4], o[:, 4] ** 2 / o[:, 5])):
            z = welch_z(gv, ov)
            zs.append(z)
            print("L=%3d T=%.4f %-2s gpu %.6f +- %.2g  oracle %.6f +- %.2g  z=%+.2f" % (
                L_, T, nm, gv.mean(), gv.std(ddof=1) / np.sqrt(len(gv)), ov.mean(),
                ov.std(ddof=1) / np.sqrt(len(ov)), z))
            assert abs(z) < 4.5, (key, nm)
    chi2, dof, lim, f2 = chi2_summary(zs)
    print("XY ladder scan: chi2 = %.1f for %d comparisons (99.9%% quantile %.1f); %.0f%% within 2 sigma" %
          (chi2, dof, lim, 100 * f2))
    assert chi2 < lim
