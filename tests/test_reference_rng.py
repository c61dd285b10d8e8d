"""A record of how the reference's random number generator biases its own
averages, and the pins of the generator the parity fixtures use instead.

The reference draws every decision from glibc rand() (isingLib.c:239-249,
heisenbergLib.c:95-112 and 441-473), an additive lagged-Fibonacci generator
(r_i = r_{i-3} + r_{i-31} up to a carry), and one Metropolis update consumes
3 (Ising) to 6 or more (XY, Heisenberg: site pair, rejection-sampled
direction, acceptance) consecutive values.  The oracle restatement with the
same algorithm and the same consumption order, but integer draws from
xoshiro256**, gives different equilibrium averages: config_fp64_glibc.json and
config_fp64.json (tests/golden/make_config_fp64.py) hold both, run on the
same points with the same seeds.  The GPU engine (Philox4x32-10) agrees with
the xoshiro column (tests/test_gpu_configs.py, test_gpu_xy_scan_validation.py),
so those are the parity references; this test keeps the evidence that the
difference is the generator and not the sampling.
"""
import json
import os

import numpy as np

from oracle import mcref

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return json.load(open(os.path.join(G, name)))


def test_xoshiro256ss_known_answers():
    """Blackman and Vigna's reference implementation from state {1, 2, 3, 4}."""
    lib = mcref.lib()
    lib.mcref_xoshiro_set(1, 2, 3, 4)
    assert [lib.mcref_xoshiro_next() for _ in range(4)] == [11520, 0, 1509978240, 1215971899390074240]


def test_rng_mode_switch_keeps_glibc_replay():
    """Switching to xoshiro and back leaves the glibc stream the replay
    fixtures use (mode 0 is rand() itself)."""
    with mcref.rng("xoshiro"):
        mcref.lib().mcref_srand(7)
        a = [mcref.lib().mcref_rand() for _ in range(3)]
    mcref.lib().mcref_srand(7)
    b = [mcref.lib().mcref_rand() for _ in range(3)]
    assert a != b and all(0 <= v < 2 ** 31 for v in a + b)
    import ctypes
    libc = ctypes.CDLL(None)
    libc.srand(7)
    assert b == [libc.rand() for _ in range(3)]


def test_glibc_rand_biases_the_reference_heisenberg_averages():
    """Triangular Neel-DMI 128^2 at (T, Bz) = (0.847, 0.323) and (1.0, 0.774):
    12 runs per generator; the energies differ by many standard errors."""
    x, g = _load("config_fp64.json")["tri_dmi128"], _load("config_fp64_glibc.json")["tri_dmi128"]
    for key in ("T26_B10", "T31_B24"):
        ex, eg = np.asarray(x["points"][key]["runs"])[:, 0], np.asarray(g["points"][key]["runs"])[:, 0]
        z = (eg.mean() - ex.mean()) / np.hypot(eg.std(ddof=1) / np.sqrt(len(eg)), ex.std(ddof=1) / np.sqrt(len(ex)))
        print(key, "E/T glibc %.6f xoshiro %.6f z=%.1f" % (eg.mean(), ex.mean(), z))
        assert abs(z) > 5.0


def test_ising_specific_heat_unbiased_under_both_generators():
    """Ising 32^2 (3 draws per update): C from both generators agrees with
    Kaufman's exact torus value at T = 2.433 and 2.833, away from T_c."""
    import exact
    for name in ("config_fp64.json", "config_fp64_glibc.json"):
        d = _load(name)["ising_square32"]
        for key in ("T7", "T10"):
            p = d["points"][key]
            T = p["T"]
            r = np.asarray(p["runs"])
            E, E2 = r[:, 4] * T, r[:, 5] * T * T       # beta units -> energy units
            c = 1024 * (E2 - E * E) / T ** 2
            _, c_ex = exact.ising_e_c(32, T)
            z = (c.mean() - c_ex) / (c.std(ddof=1) / np.sqrt(len(c)))
            print(name, key, "C %.4f exact %.4f z=%.2f" % (c.mean(), c_ex, z))
            assert abs(z) < 4.5
