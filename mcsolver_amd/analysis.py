"""Estimates with error bars from the engine's per-row accumulators and bins.

Each temperature row keeps FP64 sums over all measured frames (accumulators)
and over nbins consecutive blocks of frames (bins of e, e^2, M, M^2, M^4).
Block averages are close to independent when a block is much longer than
the autocorrelation time; the error bars are jackknife over blocks.

Definitions follow the reference (mcMain.py:251-253, win.py:136-142,
isingLib.c:433): U4 = <M^2>^2 / <M^4>, C = N (<e^2> - <e>^2) / T^2 with e the
energy per site and N the number of sites.
"""
from __future__ import annotations

import numpy as np

from . import _lib as L


def jackknife(bins: np.ndarray, fn):
    """bins [nb, q] block means; fn(mean over blocks [q]) -> scalar.
    Returns (estimate, standard error)."""
    nb = bins.shape[0]
    full = fn(bins.mean(0))
    if nb < 2:
        return full, float("nan")
    tot = bins.sum(0)
    js = np.array([fn((tot - bins[k]) / (nb - 1)) for k in range(nb)])
    err = np.sqrt((nb - 1) / nb * np.sum((js - js.mean()) ** 2))
    return full, err


def batch_estimates(batch, temps, n_sites_per_row) -> list:
    """Per-row estimates of a finished Batch run (rows in order).  The run
    must have measured exactly the planned number of frames (Batch nframes),
    so that every bin holds the same number of frames."""
    bins = batch.bin_data()
    acc = batch.accumulators()
    out = []
    for r in range(bins.shape[0]):
        cnt = acc[r, L.A_COUNT]
        nb = bins.shape[1]
        # frames are spread evenly over bins: each bin holds cnt/nb frames
        per = cnt / nb
        means = bins[r] / per
        d = {"T": float(temps[r]), "frames": int(cnt)}
        N = n_sites_per_row[r]
        T = float(temps[r])
        d["E"], d["E_err"] = jackknife(means, lambda m: m[0])
        d["C"], d["C_err"] = jackknife(means, lambda m: N * (m[1] - m[0] ** 2) / T ** 2)
        d["M"], d["M_err"] = jackknife(means, lambda m: m[2])
        d["U4"], d["U4_err"] = jackknife(means, lambda m: m[3] ** 2 / m[4])
        out.append(d)
    return out


def tau_int(x, c: float = 5.0):
    """Integrated autocorrelation time of a series, in units of its spacing:
    tau = 1/2 + sum_{t=1..W} rho(t), with rho the normalised autocovariance
    (divided by n, computed by FFT) and the window W chosen self-consistently
    as the smallest W >= c * tau(W) (Madras and Sokal, J. Stat. Phys. 50, 109,
    1988).  Returns (tau, W); W = n - 1 means the window never closed and the
    series is too short for tau.  A run of n frames holds about n / (2 tau)
    independent samples."""
    x = np.asarray(x, np.float64)
    n = x.size
    if n < 2:
        return 0.5, 0
    x = x - x.mean()
    var = float(x @ x) / n
    if var == 0.0:
        return 0.5, 0
    m = 1 << (2 * n - 1).bit_length()
    f = np.fft.rfft(x, m)
    rho = np.fft.irfft(f * np.conj(f), m)[:n] / (n * var)
    tau = 0.5
    for w in range(1, n):
        tau += float(rho[w])
        if w >= c * tau:
            return tau, w
    return tau, n - 1
