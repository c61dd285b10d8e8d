"""The closed forms themselves, and the FP64 oracle (reference algorithm)
against them."""
import numpy as np
import pytest

import exact
from mcsolver_amd import lattice as lat
from oracle import mcref


@pytest.mark.parametrize("K", [0.2, 0.4407, 0.6])
def test_kaufman_matches_brute_force_4x4(K):
    assert abs(exact.ising_logz(4, K) - exact.ising_brute_logz(4, K)) < 1e-9


def test_kaufman_energy_limits():
    e_hot, _ = exact.ising_e_c(32, 50.0)
    e_cold, _ = exact.ising_e_c(32, 0.5)
    assert abs(e_hot) < 0.05 and abs(e_cold + 2.0) < 1e-3


def test_single_spin_quadrature_limits():
    # strong field: fully polarised; no terms: <E> = 0
    e, m = exact.single_spin([0, 0, 0], 5.0, 10.0)
    assert m > 0.97
    e0, m0 = exact.single_spin([0, 0, 0], 0.0, 1.0)
    assert abs(e0) < 1e-12 and abs(m0) < 1e-12


def test_dimer_isotropic_is_langevin():
    # E = -s1.s2: <s1.s2> = coth(beta) - 1/beta
    e, _ = exact.dimer(-np.eye(3), 1.3)
    assert abs(-e - (1 / np.tanh(1.3) - 1 / 1.3)) < 1e-6


def _ring(model, L, J=-1.0):
    if model == "ising":
        return lat.build_lattice(L, 1, 1, np.eye(3), [[0, 0, 0]], [lat.Bond(0, 0, (1, 0, 0), [J])], [1.0],
                                 ising=True)
    return lat.build_lattice(L, 1, 1, np.eye(3), [[0, 0, 0]], [lat.Bond(0, 0, (1, 0, 0), [J] * 3)], [1.0])


@pytest.mark.parametrize("model,T", [("ising", 1.0), ("xy", 0.8), ("heis", 0.7)])
def test_oracle_ring_energy(model, T):
    """Random-sequential FP64 oracle on a 64-site ring vs the chain limit;
    error bar from 8 independent glibc seeds."""
    Lt = _ring(model, 64)
    es = []
    for seed in range(8):
        if model == "ising":
            a = lat.ising_args(Lt, T, 0.0, 200, 1500, algorithm="Wolff")
            out, _ = mcref.ising_run(a, seed=seed + 1)
            es.append(out[4] * T)
        else:
            a = lat.on_args(Lt, T, 0.0, 200, 1500)
            out, _, _ = mcref.on_run(2 if model == "xy" else 3, a, seed=seed + 1)
            es.append(out[8] * T)
    es = np.array(es)
    want = -exact.ring_bond_energy(model, 1.0 / T)
    assert abs(es.mean() - want) < 4 * es.std(ddof=1) / np.sqrt(len(es)) + 2e-3, (es.mean(), want)


@pytest.mark.parametrize("nc", [3, 2])
def test_oracle_single_spin_anisotropy(nc):
    """Isolated spins (no bonds) with D and h: E per site and the field-axis
    spin <s_z> (Heisenberg) or <s_x> (XY: the field acts on x and the
    reference reports the x projection, xyLib.c:203,633-654) against
    quadrature; XY with a spin length of 1.5 and D_x != D_y."""
    if nc == 3:
        D, h, T, S = (0.1, -0.2, -0.5), 0.3, 0.8, 1.0
    else:
        D, h, T, S = (0.25, -0.15, 0.0), 0.3, 0.6, 1.5
    Lt = lat.build_lattice(16, 16, 1, np.eye(3), [[0, 0, 0]], [], [S], [list(D)])
    es, ms = [], []
    for seed in range(6):
        a = lat.on_args(Lt, T, h, 100, 400)
        out, _, _ = mcref.on_run(nc, a, seed=seed + 3)
        es.append(out[8] * T)
        ms.append(out[25])
    e_ex, m_ex = exact.single_spin(D, h, 1 / T, nc=nc, S=S)
    es, ms = np.array(es), np.array(ms)
    assert abs(es.mean() - e_ex) < 4 * es.std(ddof=1) / np.sqrt(6) + 2e-3
    assert abs(ms.mean() - m_ex) < 4 * ms.std(ddof=1) / np.sqrt(6) + 2e-3
