"""The integrated autocorrelation time estimator (mcsolver_amd/analysis.py)
against closed forms."""
import numpy as np
import pytest

from mcsolver_amd.analysis import tau_int


@pytest.mark.parametrize("phi", [0.5, 0.9, 0.97])
def test_ar1_closed_form(phi):
    # x_{t+1} = phi x_t + noise: rho(t) = phi^t, tau = (1 + phi) / (2 (1 - phi))
    rng = np.random.default_rng(7)
    n = 400_000
    e = rng.normal(size=n)
    x = np.empty(n)
    x[0] = e[0] / np.sqrt(1 - phi * phi)
    for t in range(1, n):
        x[t] = phi * x[t - 1] + e[t]
    tau, w = tau_int(x)
    exact = (1 + phi) / (2 * (1 - phi))
    assert w < n - 1
    assert abs(tau - exact) < 0.06 * exact, (tau, exact)


@pytest.mark.parametrize("k", [1, 4, 16])
def test_repeated_samples(k):
    # independent values each held k frames: rho(t) = 1 - t/k for t < k, tau = k / 2
    rng = np.random.default_rng(k)
    x = np.repeat(rng.normal(size=100_000), k)
    tau, _ = tau_int(x)
    assert abs(tau - k / 2) < 0.05 * max(k / 2, 1.0), (tau, k / 2)


def test_degenerate_series():
    assert tau_int([1.0])[0] == 0.5
    assert tau_int(np.full(100, 3.0))[0] == 0.5
