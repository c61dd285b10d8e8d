"""Input-file parsing of the batched driver (fileio.py:169-292 semantics) and
the truncated dipolar bond entries (intent of Lattice.py:286-308)."""
import itertools
import os

import numpy as np
import pytest

from mcsolver_amd import driver as D
from mcsolver_amd import lattice as lat

SAMPLES = "/root/reference/samples"


def _example():
    return D.McInput(
        LMatrix=[[1.0, 0.0, 0.0], [-0.5, 0.866025404, 0.0], [0.0, 0.0, 1.0]], LPack=[6, 4, 1],
        pos=[[1 / 3, 2 / 3, 0.0], [2 / 3, 1 / 3, 0.0]], S=[1.5, 1.0], DList=[[0.1, -0.2, -0.3], [0.0, 0.0, 0.5]],
        bondList=[[0, 1, [0, 0, 0], -1.0, -1.1, -1.2, 0.1, 0.2, 0.3, -0.1, -0.2, -0.3],
                  [1, 1, [1, -1, 0], 0.5, 0.5, 0.5, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0]],
        T0=0.5, T1=2.5, nT=5, H0=-0.25, H1=0.75, nH=3, dipoleAlpha=0.0, GcOrb=[[0, 1], [1, 0, 0]],
        orbGroupList=[[0], [1]], groupInSC=True,
        localCircuitList=[((0, (0, 0, 0)), (1, (1, 0, 0)), (0, (1, 1, 0)))], spinFrame=2, nthermal=100,
        nsweep=400, ninterval=0, xAxisType="T", modelType="Heisenberg", algorithm="Metropolis", ncores=3)


def test_round_trip():
    p = _example()
    q = D.parse_input(D.format_input(p))
    for k in p.__dataclass_fields__:
        a, b = getattr(p, k), getattr(q, k)
        if k in ("LMatrix", "pos", "DList"):
            np.testing.assert_allclose(np.array(a), np.array(b), atol=1e-9)
        elif k == "bondList":
            for x, y in zip(a, b):
                assert x[:3] == y[:3]
                np.testing.assert_allclose(x[3:], y[3:], atol=1e-9)
        elif isinstance(a, float):
            assert abs(a - b) < 1e-6, k
        else:
            assert a == b, k


def test_missing_tag_and_version():
    txt = D.format_input(_example())
    with pytest.raises(ValueError, match="version"):
        D.parse_input(txt.replace("version: 3.0", "version: 2.0"))
    with pytest.raises(ValueError, match="Ncores"):
        D.parse_input(txt.replace("Ncores:", "Cores:"))


def test_positional_columns_ignore_labels():
    """Stale labels ("Jz Jx Jy", "Dz Dx Dy") do not change which number
    goes where: values are read by position."""
    txt = D.format_input(_example())
    relab = txt.replace("bond 0: Jx", "bond 0: Jz").replace("Dx 0.100000000 Dy", "Dz 0.100000000 Dx")
    q = D.parse_input(relab)
    assert q.bondList[0][3] == -1.0 and q.DList[0] == [0.1, -0.2, -0.3]


@pytest.mark.skipif(not os.path.isdir(SAMPLES), reason="reference samples not present")
def test_reference_samples():
    p = D.load_input(os.path.join(SAMPLES, "Square_XY_isotropic"))
    assert p.LPack == [16, 16, 1] and p.S == [1.0] and len(p.bondList) == 2
    assert p.bondList[1][:3] == [0, 0, [0, 1, 0]] and p.bondList[1][3:6] == [-1.0, -1.0, -1.0]
    assert (p.T0, p.T1, p.nT, p.H0, p.H1, p.nH) == (0.9, 1.2, 8, 0.0, 0.1, 1)
    assert (p.nthermal, p.nsweep, p.ninterval) == (40000, 80000, 1)
    assert (p.modelType, p.algorithm, p.ncores, p.xAxisType) == ("XY", "Wolff", 4, "T")
    assert p.orbGroupList == [[0]] and p.groupInSC and p.localCircuitList == [] and p.spinFrame == 0

    p = D.load_input(os.path.join(SAMPLES, "CrI3With2NNCoupling"))
    assert p.LMatrix[1] == [-0.5, 0.866025404, 0.0] and p.S == [1.5, 1.5]
    # labelled "Dz Dx Dy" in the file; read positionally as Dx Dy Dz
    assert p.DList[0] == [-3.12276251875, 0.0, 0.0]
    assert len(p.bondList) == 12 and p.bondList[11][:3] == [0, 1, [-1, -1, 0]]
    assert p.bondList[0][3:6] == [-19.49182553875, -18.47237479, -18.47237479]
    assert (p.T0, p.T1, p.nT) == (30.0, 50.0, 21) and p.orbGroupList == [[0], [1]]
    assert (p.nthermal, p.nsweep, p.ninterval, p.modelType, p.algorithm) == (80000, 640000, 0, "Heisenberg",
                                                                             "Metropolis")

    p = D.load_input(os.path.join(SAMPLES, "SkyrmionOnSqaureLattice"))
    assert p.bondList[0][3:] == [-1.0, -1.0, -1.0, 0.0, 1.0, 0.0, 0.0, -1.0, 0.0]
    assert p.localCircuitList == [((0, (0, 0, 0)), (0, (1, 0, 0)), (0, (1, 1, 0))),
                                  ((0, (0, 0, 0)), (0, (1, 1, 0)), (0, (0, 1, 0)))]
    assert (p.H0, p.H1, p.nH, p.xAxisType, p.DList) == (0.0, 0.7, 8, "H", [[0.0, 0.0, -0.1]])

    p = D.load_input(os.path.join(SAMPLES, "SkyrmionOnHexLattice"))
    assert len(p.localCircuitList) == 4 and p.spinFrame == 1 and len(p.pos) == 2
    assert p.localCircuitList[0] == ((1, (0, 0, 0)), (1, (1, 0, 0)), (0, (1, 0, 0)))


def test_input_lattice_matches_builder():
    p = _example()
    Lt = D.build_input_lattice(p)
    ref = lat.build_lattice(6, 4, 1, p.LMatrix, p.pos, [lat.Bond(b[0], b[1], tuple(b[2]), b[3:]) for b in p.bondList],
                            p.S, p.DList, ki_s=0, ki_t=1, ki_overLat=(1, 0, 0), orbGroupList=[[0], [1]],
                            groupInSC=True, localCircuitList=p.localCircuitList)
    np.testing.assert_array_equal(Lt.nbr, ref.nbr)
    np.testing.assert_array_equal(Lt.J, ref.J)
    np.testing.assert_array_equal(Lt.circuits, ref.circuits)
    assert Lt.groups == ref.groups


# ---------------------------------------------------------------------------
# dipolar couplings
# ---------------------------------------------------------------------------

def _direct_dipole(pos, s, alpha, cutoff=None):
    """sum over unordered pairs of alpha/r^3 (s_i.s_j - 3 (s_i.u)(s_j.u))."""
    E = 0.0
    for i, j in itertools.combinations(range(len(pos)), 2):
        r = pos[j] - pos[i]
        d = np.linalg.norm(r)
        if cutoff is not None and d > cutoff:
            continue
        u = r / d
        E += alpha / d ** 3 * (s[i] @ s[j] - 3 * (s[i] @ u) * (s[j] @ u))
    return E


@pytest.mark.parametrize("cutoff", [None, 1.6, 2.05])
def test_dipolar_energy_direct_sum(cutoff):
    base = lat.build_lattice(3, 3, 2, [[1, 0, 0], [-0.5, 0.8660254, 0], [0, 0, 1.3]], [[0, 0, 0], [0.5, 0.3, 0.1]],
                             [lat.Bond(0, 1, (0, 0, 0), [-1.0, -1.0, -1.0])], [1.0, 1.0])
    Lt = lat.dipolar_bonds(base, 0.37, cutoff=cutoff)
    rng = np.random.default_rng(3)
    s = rng.normal(size=(Lt.n, 3))
    s /= np.linalg.norm(s, axis=1, keepdims=True)
    # only the appended links carry dipoles: energy of those alone
    E_tab = 0.0
    for i in range(Lt.n):
        for l in range(base.nlink[i], Lt.nlink[i]):
            E_tab += 0.5 * s[i] @ lat.tensor9_to_mat(Lt.J[i, l]) @ s[Lt.nbr[i, l]]
    E_dir = _direct_dipole(Lt.pos, s, 0.37, cutoff)
    assert abs(E_tab - E_dir) < 1e-10 * max(1.0, abs(E_dir))
    # the original links are untouched
    np.testing.assert_array_equal(Lt.nbr[:, :base.maxl], base.nbr)
    # every link i->j has the mirror j->i with the transposed (here equal) tensor
    for i in range(Lt.n):
        for l in range(base.nlink[i], Lt.nlink[i]):
            j = Lt.nbr[i, l]
            ls = [m for m in range(base.nlink[j], Lt.nlink[j]) if Lt.nbr[j, m] == i]
            assert len(ls) == 1
            np.testing.assert_allclose(lat.tensor9_to_mat(Lt.J[j, ls[0]]), lat.tensor9_to_mat(Lt.J[i, l]).T)


def test_dipolar_ising_is_isotropic_part():
    base = lat.square(4, J=-1.0, ising=True)
    Lt = lat.dipolar_bonds(base, 0.5, ising=True)
    sg = np.where(np.random.default_rng(0).random(Lt.n) < 0.5, -1.0, 1.0)
    E_tab = sum(0.5 * Lt.J[i, l, 0] * sg[i] * sg[Lt.nbr[i, l]] for i in range(Lt.n)
                for l in range(base.nlink[i], Lt.nlink[i]))
    E_dir = sum(0.5 / np.linalg.norm(Lt.pos[i] - Lt.pos[j]) ** 3 * sg[i] * sg[j]
                for i, j in itertools.combinations(range(Lt.n), 2))
    assert abs(E_tab - E_dir) < 1e-10
    assert Lt.nlink.min() == base.nlink.min() + Lt.n - 1


def test_dipolar_link_limit():
    base = lat.square(8, J=-1.0)
    with pytest.raises(ValueError, match="cutoff"):
        lat.dipolar_bonds(base, 0.1, max_links=1000)


# ---------------------------------------------------------------------------
# output formats (mcMain.py:126-135, 263-265, 274-323; win.py:151-156)
# ---------------------------------------------------------------------------

def test_ising_output_lines(tmp_path):
    Lt = lat.square(2, J=-1.0, ising=True)
    data = (0.5, 0.25, 0.125, 0.01, -1.5, 2.5, -0.75, 0.625, 0.6, 0.375, ((1.0, -1.0, 1.0, 1.0),))
    r = D.PointResult(0, 0.05, 0.2, 0.1, data)     # T clamped to 0.1
    w = D._Writer(str(tmp_path), verbose=False)
    w.headers()
    si, sj, sij, auto, E, E2, U4 = w.ising(Lt, r, 1)
    assert (E, E2) == (-1.5 * 0.1, 2.5 * 0.1 ** 2)
    out = (tmp_path / "out").read_text().split("\n")
    assert out[0] == "#T #H"
    assert out[1] == "0.100 0.200 0.500 0.250 0.375 0.125 -0.150 0.025 -0.075 0.006 0.600000"
    fr = (tmp_path / "IsingSpinDistribution.T0.100.H0.200.0.txt").read_text().split("\n")
    assert fr[0] == "#x       #y       #z       #spin"
    assert fr[2] == "0.000000 1.000000 0.000000 -1.000\n".strip()
    w.result([(0.05, 0.2, si, sj, (sij - si * sj) / 0.05, E, (E2 - E * E) / 0.05 ** 2 * 4, 0.0, U4, auto)])
    res = (tmp_path / "result.txt").read_text().split("\n")
    assert res[0].startswith("#Temp          #Field         #<Si>")
    # Susc = (0.125 - 0.5*0.25)/0.05 = 0; C = (0.025 - 0.0225)/0.05^2 * 4 = 4
    assert res[1] == ("   5.000000E-02   2.000000E-01   5.000000E-01   2.500000E-01   0.000000E+00  -1.500000E-01"
                      "   4.000000E+00   0.000000E+00   6.000000E-01   1.000000E-02")


def test_on_output_lines(tmp_path):
    Lt = lat.build_lattice(1, 1, 1, np.eye(3), [[0, 0, 0], [0.5, 0, 0]], [lat.Bond(0, 1, (0, 0, 0), [-1, -1, -1])],
                           [1.0, 1.0], orbGroupList=[[0], [1]], groupInSC=True)
    vals = [0.6, 0.0, 0.8, 0.0, 0.6, 0.8, 0.9, 0.02, -2.0, 4.5, 0.7, 0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7, -1.0, 1.5,
            0.11, 0.22, 0.33, 0.44, 0.55, 0.66, 0.99]
    frames = (((1.0, 0.0, 0.0), (0.0, 1.0, 0.0)),)
    groups = tuple(float(k) for k in range(12))
    r = D.PointResult(0, 2.0, 0.5, 2.0, tuple(vals) + (frames, groups))
    w = D._Writer(str(tmp_path), verbose=False)
    w.headers()
    si, sj, sij, auto, E, E2, U4, Q = w.on(Lt, r, 1)
    line = (tmp_path / "out").read_text().split("\n")[1]
    assert line.startswith("T= 2.000000E+00 h= 5.000000E-01 <Siz>= 1.100000E-01 <Sjz>= 2.200000E-01 ")
    assert "<Si>= 1.000000E+00 <Sj>= 1.000000E+00 <SiSj>= 9.000000E-01 <Si><Sj>= 6.400000E-01 " in line
    assert "<E>= -4.000000E+00 <E^2>= 1.800000E+01 <U4>= 7.000000E-01 " in line
    # spin_i_r = (0.1, 0.2, 0.3), spin_j_r = (0.4, 0.5, 0.6): dot 0.32; spin_ij_r = 0.7
    assert "<SiSj>r= 7.000000E-01 <Si>r<Sj>r= 3.200000E-01 <Er>= -2.000000E+00 <E^2_r>= 6.000000E+00 " in line
    assert line.endswith("<Q>= 9.900000E-01")
    sds = (tmp_path / "spinDotSpin.txt").read_text().split("\n")
    assert sds[1] == "2.000 0.500 " + "".join("%.6f " % k for k in range(12))
    fr = (tmp_path / "OnSpinDistribution.T2.000.H0.500.0.txt").read_text().split("\n")
    assert fr[2] == "0.500000 0.000000 0.000000 0.000000 1.000000 0.000000"
