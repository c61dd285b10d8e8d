"""The exact Ising acceptance table (engine.exact_ising_table), on the CPU:
when it applies, its unit, its coverage of every possible dE, and its
thresholds against floor(2^32 exp(-beta dE)); the checker follows it."""
import itertools

import numpy as np

from mcsolver_amd import lattice as lat
from mcsolver_amd.engine import compile_graph, exact_ising_table
from oracle import mcref


def _graph(Lt):
    return compile_graph(1, Lt.nbr, Lt.J[:, :, :1], Lt.S, Lt.D)


def test_applies_only_to_dyadic_inputs():
    g = _graph(lat.square(8, ising=True))
    assert exact_ising_table([g], [0.4], [0.0]) is not None
    assert exact_ising_table([g], [0.4], [0.25]) is not None
    assert exact_ising_table([g], [0.4], [0.1]) is None                       # 0.1 is no multiple of 2^-k
    assert exact_ising_table([_graph(lat.square(8, J=-19.49, ising=True))], [0.4], [0.0]) is None
    assert exact_ising_table([_graph(lat.square(8, S=1.5, ising=True))], [0.4], [0.0]) is not None


def test_unit_and_thresholds():
    g = _graph(lat.square(8, ising=True))                                    # J S_t = 1, S = 1, z = 4
    betas = [0.2, 0.4406868, 1.0]
    thr, inv_unit = exact_ising_table([g], betas, [0.0])
    assert inv_unit == 0.5                                                    # dE is a multiple of 2
    assert thr.shape == (3, 5)                                                # |dE| <= 2 * 4 = 8 = 4 units
    for r, b in enumerate(np.asarray(betas, np.float32).astype(np.float64)):
        for k in range(thr.shape[1]):
            want = min(np.floor(np.exp(-b * 2.0 * k) * 2.0 ** 32), 2.0 ** 32 - 1)
            assert thr[r, k] == want


def test_table_covers_every_dE():
    """Two orbitals, spin lengths 1 and 1.5, unequal dyadic couplings and a
    field: every dE = -2 s_i (f - h) over all neighbour configurations of every
    site is an exact multiple of the unit below the table's length."""
    bonds = [lat.Bond(0, 1, (0, 0, 0), [-1.0]), lat.Bond(1, 0, (1, 0, 0), [-0.75]), lat.Bond(0, 0, (0, 1, 0), [0.5]),
             lat.Bond(1, 1, (0, 1, 0), [-1.25])]
    Lt = lat.build_lattice(4, 4, 1, np.eye(3), [[0, 0, 0], [0.5, 0, 0]], bonds, [1.0, 1.5], ising=True)
    g = _graph(Lt)
    h = 0.375
    thr, inv_unit = exact_ising_table([g], [0.5], [h])
    cj = g.coup.reshape(-1, 9)[:, 0].astype(np.float32)
    S = g.clsdat.reshape(-1, 8)[:, 0].astype(np.float32)
    link = g.link.reshape(g.n, -1)
    for p in range(g.n):
        c = cj[link[p] >> 24]
        for sig in itertools.product((-1.0, 1.0), repeat=len(c) + 1):
            f = np.float32(0.0)
            for cl, sj in zip(c, sig[1:]):
                f = np.float32(f + np.float32(cl) * np.float32(sj))
            s = np.float32(sig[0]) * S[g.cls[p]]
            dE = np.float32(np.float32(-2.0) * s) * np.float32(f - np.float32(h))
            k = float(dE) * inv_unit
            assert k == round(k) and abs(k) < thr.shape[1]


def test_checker_uses_the_table():
    """With all thresholds 0 only flips with dE <= 0 are accepted, so the
    energy never rises; with the real table it does at high temperature."""
    Lt = lat.square(8, ising=True)
    g = _graph(Lt)
    rng = np.random.default_rng(1)
    s0 = np.where(rng.random((1, 1, Lt.n)) < 0.5, 1.0, -1.0).astype(np.float32)
    thr, iu = exact_ising_table([g], [0.1], [0.0])
    nbr, J, S, D = Lt.nbr, Lt.J[:, :, :1], Lt.S, Lt.D
    e = []
    for tab in ((np.zeros_like(thr), iu), (thr, iu)):
        sp = s0.copy()
        es = [mcref.cb_energy(1, nbr, J, S, D, 0.0, sp[0])[0]]
        for k in range(20):
            mcref.cb_sweeps(1, 2, nbr, J, S, D, g.colour, [0.1], [0.0], np.array([0], np.uint32), 7, k, 1, sp,
                            ising_table=tab)
            es.append(mcref.cb_energy(1, nbr, J, S, D, 0.0, sp[0])[0])
        e.append(np.array(es))
    assert np.all(np.diff(e[0]) <= 0)
    assert np.any(np.diff(e[1]) > 0)
