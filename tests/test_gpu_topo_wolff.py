"""Topological charge of a known texture, and Wolff against closed forms."""
import numpy as np
import pytest

import exact
from mcsolver_amd import _lib as L
from mcsolver_amd import lattice as lat
from mcsolver_amd.engine import Batch, ReplicaSpec, compile_graph
from mcsolver_amd.workloads import triangular_neel_dmi

pytestmark = pytest.mark.gpu


def _skyrmion(Lt, R0=8.0, core_down=True):
    pos = Lt.pos
    c = pos.mean(0)
    d = pos[:, :2] - c[:2]
    r = np.linalg.norm(d, axis=1)
    phi = np.arctan2(d[:, 1], d[:, 0])
    theta = np.pi * np.exp(-(r / R0) ** 2)  # pi at the core, 0 far away
    if not core_down:
        theta = np.pi - theta
    return np.stack([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)], 1)


def _q_numpy(Lt, s):
    """Berg-Luscher solid angles with the reference's atan form
    (heisenbergLib.c:114-127) over the lattice's circuits."""
    q = 0.0
    for a, b, c in Lt.circuits:
        s1, s2, s3 = s[a], s[b], s[c]
        re = 1 + s1 @ s2 + s2 @ s3 + s3 @ s1
        im = s1 @ np.cross(s2, s3)
        q += (np.pi if im > 0 else -np.pi) if abs(re) < 1e-6 else 2 * np.arctan(im / re)
    return q / 3.1415926535 / 4


@pytest.mark.parametrize("texture", ["skyrmion", "antiskyrmion_field", "uniform"])
def test_topological_charge(gpu, texture):
    Lt = triangular_neel_dmi(48)
    if texture == "uniform":
        s = np.zeros((Lt.n, 3)); s[:, 2] = 1
    else:
        s = _skyrmion(Lt, core_down=(texture == "skyrmion"))
    g = compile_graph(3, Lt.nbr, Lt.J, Lt.S, Lt.D, circuits=Lt.circuits)
    b = Batch(3, [g], [ReplicaSpec(0, 0)], [1.0], [0.0], [s], meas=L.MEAS_Q)
    b.init()
    b.measure(0)
    q = b.accumulators()[0, L.A_Q]
    want = _q_numpy(Lt, s)
    print(texture, q, want)
    assert abs(q - want) < 1e-4
    if texture == "uniform":
        assert abs(q) < 1e-6
    else:
        assert abs(abs(q) - 1.0) < 0.02


@pytest.mark.parametrize("kind", ["ising", "xy", "heis"])
def test_wolff_physics(gpu, kind):
    if kind == "ising":
        Lt = lat.square(16, J=-1.0, ising=True)
        nc, T = 1, 2.269
    else:
        Lt = lat.build_lattice(128, 1, 1, np.eye(3), [[0, 0, 0]], [lat.Bond(0, 0, (1, 0, 0), [-1.0] * 3)], [1.0])
        nc, T = (2, 0.8) if kind == "xy" else (3, 0.7)
    J = Lt.J if nc > 1 else Lt.J[:, :, :1]
    g = compile_graph(nc, Lt.nbr, J, Lt.S, Lt.D)
    R = 8
    rng = np.random.default_rng(1)
    init = [np.ones(Lt.n) if nc == 1 else rng.normal(size=(Lt.n, nc)) for _ in range(R)]
    init = [v if nc == 1 else v / np.linalg.norm(v, axis=1, keepdims=True) for v in init]
    b = Batch(nc, [g], [ReplicaSpec(0, k) for k in range(R)], [1 / T] * R, [0.0] * R, init, nbins=4, nframes=3000,
              wolff=True)
    b.init()
    b.run(100, 3000, 1, algorithm=L.WOLFF, clusters_per_sweep=16)
    A = b.accumulators()
    e = A[:, L.A_E] / A[:, L.A_COUNT]
    if kind == "ising":
        want, _ = exact.ising_e_c(16, T)
    else:
        want = -exact.ring_bond_energy(kind, 1 / T)
    print(kind, e.mean(), e.std(ddof=1) / np.sqrt(R), want)
    assert abs(e.mean() - want) < 4 * e.std(ddof=1) / np.sqrt(R) + 5e-4
