"""Replay cases for the reference engines: lattice specs and run parameters.

Shared by make_fixtures.py (which runs the reference's own compiled C, in the
container where /root/reference exists) and tests/test_oracle_replay.py
(which runs our FP64 restatement on the same inputs and compares).
"""
from __future__ import annotations

import hashlib

import numpy as np

from mcsolver_amd import lattice as lat


def _honeycomb(L, seed, S=1.5, Dz=-0.1):
    rng = np.random.default_rng(seed)
    bonds = []
    for (s, t, o) in [(0, 1, (0, 0, 0)), (0, 1, (-1, 0, 0)), (0, 1, (0, 1, 0)), (0, 0, (1, 0, 0)), (0, 0, (1, 1, 0)),
                      (0, 0, (0, 1, 0)), (1, 1, (1, 0, 0)), (1, 1, (1, 1, 0)), (1, 1, (0, 1, 0))]:
        J = np.zeros(9)
        J[:3] = -1 + rng.uniform(-0.1, 0.1, 3)
        J[3:] = rng.uniform(-0.05, 0.05, 6)
        bonds.append(lat.Bond(s, t, o, J))
    return lat.build_lattice(L, L, 1, [[1, 0, 0], [-0.5, np.sqrt(3) / 2, 0], [0, 0, 1]],
                             [[1 / 3, 2 / 3, 0], [2 / 3, 1 / 3, 0]], bonds, [S, S], [[0, 0, Dz], [0, 0, Dz]],
                             localCircuitList=[((0, (0, 0, 0)), (1, (0, 0, 0)), (0, (1, 0, 0)))],
                             orbGroupList=[[0], [1]], groupInSC=True)


def _triangular_dmi(L, D0=0.3, Dz=-0.05):
    """Triangular lattice, J = -1 (ferro in the reference's sign), Neel DMI
    D = -D0 z x r_ij on each bond, easy axis."""
    a1, a2 = np.array([1.0, 0, 0]), np.array([0.5, np.sqrt(3) / 2, 0])
    bonds = []
    for o in [(1, 0, 0), (0, 1, 0), (-1, 1, 0)]:
        r = o[0] * a1 + o[1] * a2
        r = r / np.linalg.norm(r)
        Dv = -D0 * np.cross([0, 0, 1.0], r)
        M = -np.eye(3)
        # s_i^T A s_j = D . (s_i x s_j)  with A_ab = eps_{c a b} D_c
        A = np.array([[0, Dv[2], -Dv[1]], [-Dv[2], 0, Dv[0]], [Dv[1], -Dv[0], 0]])
        bonds.append(lat.Bond(0, 0, o, lat.mat_to_tensor9(M + A)))
    return lat.build_lattice(L, L, 1, [a1, a2, [0, 0, 1]], [[0, 0, 0]], bonds, [1.0], [[0, 0, Dz]],
                             localCircuitList=[((0, (0, 0, 0)), (0, (1, 0, 0)), (0, (0, 1, 0))),
                                               ((0, (1, 0, 0)), (0, (1, 1, 0)), (0, (0, 1, 0)))])


def _ising_aniso(L):
    bonds = [lat.Bond(0, 0, (1, 0, 0), [-1.0]), lat.Bond(0, 0, (0, 1, 0), [-0.4])]
    return lat.build_lattice(L, L, 1, np.eye(3), [[0, 0, 0]], bonds, [1.0], ising=True)


def _ising_2orb(L):
    bonds = [lat.Bond(0, 1, (0, 0, 0), [-1.0]), lat.Bond(1, 0, (1, 0, 0), [-0.7]), lat.Bond(0, 0, (0, 1, 0), [0.3]),
             lat.Bond(1, 1, (0, 1, 0), [-1.2])]
    return lat.build_lattice(L, L, 1, np.eye(3), [[0, 0, 0], [0.5, 0, 0]], bonds, [1.0, 1.5], ising=True,
                             ki_s=0, ki_t=1)


# name -> (model, lattice factory, args factory kwargs, srand seed)
CASES = {
    "ising_sq6_metropolis": ("ising", lambda: lat.square(6, ising=True),
                             dict(T=2.5, h=0.1, nthermal=20, nsweep=50, algorithm="Metropolis", spinFrame=2), 7),
    "ising_sq6_wolff": ("ising", lambda: lat.square(6, ising=True),
                        dict(T=2.2, h=0.05, nthermal=10, nsweep=30, algorithm="Wolff", spinFrame=0), 8),
    "ising_aniso_quirk": ("ising", lambda: _ising_aniso(6),
                          dict(T=1.9, h=0.0, nthermal=15, nsweep=40, algorithm="Metropolis", spinFrame=0), 9),
    "ising_2orb_ninterval": ("ising", lambda: _ising_2orb(6),
                             dict(T=3.0, h=0.2, nthermal=30, nsweep=60, ninterval=17, algorithm="Metropolis",
                                  spinFrame=3), 10),
    "ising_2orb_wolff": ("ising", lambda: _ising_2orb(6),
                         dict(T=2.0, h=0.3, nthermal=10, nsweep=20, ninterval=20, algorithm="Wolff", spinFrame=0),
                         12),
    "heis_honeycomb_metropolis": ("heis", lambda: _honeycomb(6, 1),
                                  dict(T=1.2, h=0.3, nthermal=10, nsweep=30, algorithm="Metropolis", spinFrame=1),
                                  11),
    "heis_honeycomb_wolff": ("heis", lambda: _honeycomb(6, 2),
                             dict(T=0.9, h=0.1, nthermal=4, nsweep=10, algorithm="Wolff", spinFrame=0), 13),
    "heis_triangular_dmi": ("heis", lambda: _triangular_dmi(6),
                            dict(T=0.5, h=0.4, nthermal=10, nsweep=30, algorithm="Metropolis", spinFrame=0), 14),
    "heis_sq6_diag_flunc": ("heis", lambda: lat.square(6, J=-1.0, S=1.0, D=(0.05, 0.0, -0.2)),
                            dict(T=1.5, h=0.2, nthermal=10, nsweep=30, algorithm="Metropolis", spinFrame=0,
                                 flunc=0.3), 15),
}


def case_args(name: str) -> tuple:
    model, mk, kw, _ = CASES[name]
    L = mk()
    kw = dict(kw)
    T, h = kw.pop("T"), kw.pop("h")
    if model == "ising":
        return lat.ising_args(L, T, h, **kw)
    return lat.on_args(L, T, h, **kw)


def args_digest(args: tuple) -> str:
    """SHA-256 of the positional arguments (everything but the callback)."""
    m = hashlib.sha256()
    for a in args[:-1]:
        if isinstance(a, tuple):
            m.update(np.asarray(a, dtype=np.float64).tobytes())
        else:
            m.update(repr(a).encode())
    return m.hexdigest()
