"""FP32 and FP16 spin storage against the FP64 oracle on the CrI3-like
honeycomb lattice (64x64 cells, full-tensor bonds, Dz = -0.1): the bias of
E, C, |M|, chi and U4 at four temperatures of the honeycomb ladder, in units
of the combined standard error of R independent runs per side.

The FP64 side is tests/golden/honeycomb_fp64.json (oracle restatement of
heisenberglib.MCMainFunction with integer draws from xoshiro256**, 32 runs per
temperature, written by tests/golden/make_honeycomb_fp64.py; with glibc
rand(), as the reference draws, its C at T = 1.005 sits 1.5 standard errors
higher).  The FP32 bias must also pass a chi-square over its 20
comparisons; FP16 is reported.
The GPU side runs the drop-in heisenberglib path (mcsolver_amd.dropin.on_main)
with the same arguments.  Observables per run, from the reference's tuple
(tuple[8] E, tuple[9] E^2 in beta units, tuple[22] |M| per cell) and the
total-spin group moments <M^2>, <M^4>:
  C = nLat (E2 - E^2),  chi = nLat (M2 - |M|^2) / T,  U4 = M2^2 / M4.
Each bias must lie within 4 combined standard errors plus a floor for the
storage precision: 1e-4 relative for FP32, 4e-3 relative for FP16 (spins
rounded to 11 significant bits after every accepted update).
"""
import json
import os

import numpy as np
import pytest

from stats import chi2_summary, welch_z

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "honeycomb_fp64.json")
NAMES = ("E", "C", "M", "chi", "U4")


def _obs(m, T, nlat):
    E, E2, M, M2, M4 = np.asarray(m, float).T
    return np.stack([E, nlat * (E2 - E * E), M, nlat * (M2 - M * M) / T, M2 * M2 / M4], axis=1)


def test_fp32_fp16_bias_vs_fp64_oracle(gpu):
    from make_honeycomb_fp64 import lattice
    from mcsolver_amd import lattice as lat
    from mcsolver_amd.dropin import on_main

    d = json.load(open(FIX))
    H = lattice()
    nlat = d["nLat"]
    assert H.n == d["n"]
    R = d["runs_per_T"]
    rows, bad = [], []
    zs = {"f32": [], "f16": []}
    for i, T in enumerate(d["temps"]):
        ref = _obs(d["runs"]["%.6f" % T], T, nlat)
        a = lat.on_args(H, T, 0.0, d["nthermal"], d["nsweep"])
        for store, floor in (("f32", 1e-4), ("f16", 4e-3)):
            runs = []
            for r in range(R):
                out = on_main(3, *a, seed=77000 + 1000 * i + r, store=store)
                g = out[28]
                runs.append([out[8], out[9], out[22], g[3], g[5]])
            got = _obs(runs, T, nlat)
            print("T=%.3f %s: %d runs done" % (T, store, R), flush=True)
            for k, nm in enumerate(NAMES):
                mg, mr = got[:, k].mean(), ref[:, k].mean()
                se = np.sqrt(got[:, k].var(ddof=1) / R + ref[:, k].var(ddof=1) / R)
                z = (mg - mr) / se if se > 0 else 0.0
                zs[store].append(welch_z(got[:, k], ref[:, k]))
                rows.append("T=%.3f %-3s %-3s gpu %.6g  fp64 %.6g  bias %+.3g (%+.2f sigma)" %
                            (T, store, nm, mg, mr, mg - mr, z))
                if abs(mg - mr) > 4 * se + floor * abs(mr):
                    bad.append(rows[-1])
    print("\n".join(rows), flush=True)
    for store in ("f32", "f16"):
        chi2, dof, lim, f2 = chi2_summary(zs[store])
        print("%s: chi2 = %.1f for %d comparisons (99.9%% quantile %.1f); %.0f%% within 2 sigma" %
              (store, chi2, dof, lim, 100 * f2))
        if store == "f32":
            assert chi2 < lim, "FP32 storage: chi2 %.1f >= %.1f" % (chi2, lim)
    assert not bad, "\n".join(bad)
