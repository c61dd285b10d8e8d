"""Statistical comparison of two sets of independent run means.

welch_z: the difference of the means of the GPU runs and the oracle runs in
units of their combined standard error (Welch), converted to the standard
normal deviate of the same two-sided p-value, so that the z of a few runs
(Student-t tails) and of many runs can be summed in one chi-square.
A relative floor (storage precision) and an absolute floor are added to the
standard error in quadrature.

chi2_summary: sum of z^2 over a test's comparisons, its degrees of freedom
and the 99.9 % quantile it must stay below.
"""
from __future__ import annotations

import numpy as np
from scipy import stats


def welch_z(g, o, rel_floor: float = 0.0, abs_floor: float = 0.0) -> float:
    g = np.asarray(g, float)
    o = np.asarray(o, float)
    vg, vo = g.var(ddof=1) / len(g), o.var(ddof=1) / len(o)
    fl = (rel_floor * abs(o.mean())) ** 2 + abs_floor ** 2
    se2 = vg + vo + fl
    d = g.mean() - o.mean()
    if se2 <= 0:
        return 0.0 if d == 0 else float("inf")
    t = d / np.sqrt(se2)
    # Welch-Satterthwaite degrees of freedom (the floor counts as known)
    den = (vg * vg / (len(g) - 1) if vg > 0 else 0.0) + (vo * vo / (len(o) - 1) if vo > 0 else 0.0)
    df = (vg + vo) ** 2 / den if den > 0 else 1e9
    df = max(df, 1.0)
    p = 2.0 * stats.t.sf(abs(t), df)
    if p <= 0:
        return float(np.sign(t)) * 40.0
    return float(np.sign(t) * stats.norm.isf(p / 2.0))


def chi2_summary(zs, q: float = 0.999):
    z = np.asarray(zs, float)
    chi2 = float(np.sum(z * z))
    dof = len(z)
    return chi2, dof, float(stats.chi2.ppf(q, dof)), float(np.mean(np.abs(z) < 2.0))
