"""The table compiler: device tables decode back to the per-site links."""
import numpy as np
import pytest

from mcsolver_amd import lattice as lat
from mcsolver_amd.engine import compile_graph

from cases import _honeycomb, _triangular_dmi

ROWMAJOR = [0, 3, 4, 6, 1, 5, 7, 8, 2]


def _check(nc, L_):
    J = L_.J if nc > 1 else L_.J[:, :, :1]
    g = compile_graph(nc, L_.nbr, J, L_.S, L_.D)
    n, z = L_.nbr.shape
    # tables are in colour-order positions: position p is site col_site[p]
    assert sorted(g.col_site.tolist()) == list(range(n))
    assert np.array_equal(g.pos[g.col_site], np.arange(n))
    linkp = g.link.reshape(n, z).astype(np.int64)
    link = np.empty_like(linkp)
    link[g.col_site] = linkp
    nbr = g.col_site[link & 0xFFFFFF]
    cid = link >> 24
    coup = g.coup.reshape(-1, 9)
    valid = L_.nbr >= 0
    assert np.array_equal(nbr[valid], L_.nbr[valid])
    if (~valid).any():
        assert (nbr[~valid] == np.nonzero(~valid)[0]).all()
    assert np.all(coup[0] == 0)
    for i in range(n):
        for l in range(z):
            if not valid[i, l]:
                assert cid[i, l] == 0
                continue
            if nc == 1:
                assert coup[cid[i, l], 0] == np.float32(L_.J[i, l, 0] * L_.S[L_.nbr[i, l]])
                assert coup[cid[i, l], 1] == np.float32(L_.J[i, l, 0])
            else:
                want = L_.J[i, l, ROWMAJOR].astype(np.float32)
                if nc == 2:  # XY keeps xx, xy, yx, yy (xyLib.c:150-153)
                    want[[2, 5, 6, 7, 8]] = 0
                assert np.array_equal(coup[cid[i, l]], want)
    # proper colouring
    col = g.colour
    for i in range(n):
        for l in range(z):
            j = L_.nbr[i, l]
            if j >= 0 and j != i:
                assert col[i] != col[j]
    assert sorted(g.col_site.tolist()) == list(range(n))
    assert g.col_ptr[-1] == n
    # class table (cls is per position)
    cd = g.clsdat.reshape(-1, 8)
    cls = np.empty_like(g.cls)
    cls[g.col_site] = g.cls
    assert np.array_equal(cd[cls, 0], L_.S.astype(np.float32))
    assert np.array_equal(cd[cls, 1:4], L_.D.astype(np.float32))
    s32 = L_.S.astype(np.float32)
    assert np.array_equal(cd[cls, 4], np.float32(1) / (s32 * s32))
    # colour classes are contiguous position ranges
    for c in range(g.ncol):
        assert (col[g.col_site[g.col_ptr[c]:g.col_ptr[c + 1]]] == c).all()
    # link patterns in colour-permuted positions
    if g.npat:
        pt = g.pattab.view(np.uint32).reshape(g.npat, z + 1).astype(np.int64)
        for p in range(n):
            i = g.col_site[p]
            row = pt[g.pat[p]]
            d = ((row[:z] & 0xFFFFFF) ^ 0x800000) - 0x800000
            assert np.array_equal(g.col_site[p + d], nbr[i]), p
            assert np.array_equal(row[:z] >> 24, cid[i])
            assert row[z] == cls[i]
    return g


@pytest.mark.parametrize("mk", [lambda: (1, lat.square(8, ising=True)), lambda: (2, lat.square(6, J=-1.0)),
                                lambda: (3, _honeycomb(4, 1)), lambda: (3, _triangular_dmi(6)),
                                lambda: (3, lat.cubic(4))])
def test_tables_roundtrip(mk):
    nc, L_ = mk()
    g = _check(nc, L_)
    assert g.npat > 0 and g.npat <= 64


def test_square_lattice_is_checkerboard():
    g = compile_graph(2, lat.square(8).nbr, lat.square(8).J, np.ones(64), np.zeros((64, 3)))
    assert g.ncol == 2
    x, y = np.divmod(np.arange(64), 8)
    assert len(set(zip(g.colour, (x + y) % 2))) == 2


def test_random_graph_has_no_patterns():
    rng = np.random.default_rng(0)
    n, z = 400, 4
    nbr = rng.integers(0, n, (n, z))
    g = compile_graph(3, nbr, -np.ones((n, z, 9)), np.ones(n), np.zeros((n, 3)))
    assert g.npat == 0


def test_bad_indices_rejected():
    with pytest.raises(ValueError):
        compile_graph(3, np.array([[1], [5]]), np.zeros((2, 1, 9)), np.ones(2), np.zeros((2, 3)))


def test_wide_coupling_tables():
    """More than 256 distinct tensors: per-link coupling indices."""
    rng = np.random.default_rng(3)
    n, z = 200, 4
    nbr = rng.integers(0, n, (n, z))
    J = rng.normal(size=(n, z, 9))
    g = compile_graph(3, nbr, J, np.ones(n), np.zeros((n, 3)))
    assert g.lcid is not None and g.npat == 0
    coup = g.coup.reshape(-1, 9)
    cid = np.empty((n, z), np.int64)
    cid[g.col_site] = g.lcid.reshape(n, z)
    lk = np.empty((n, z), np.int64)
    lk[g.col_site] = g.link.reshape(n, z).astype(np.int64)
    assert np.array_equal(g.col_site[lk], nbr)
    for i in range(n):
        for l in range(z):
            assert np.array_equal(coup[cid[i, l]], J[i, l, ROWMAJOR].astype(np.float32))


def test_uniform_coupling_flag():
    """ucid names the one tensor of a graph whose sites all carry zmax links
    with the same coupling (the Metropolis field is then J times the sum of
    the neighbour spins); any other graph has ucid -1."""
    sq = lat.square(8, J=-1.0)
    g = compile_graph(2, sq.nbr, sq.J, sq.S, sq.D)
    assert g.ucid >= 0
    assert np.array_equal(g.coup.reshape(-1, 9)[g.ucid], g.coup.reshape(-1, 9)[int(g.link[0] >> 24)])
    tri = _triangular_dmi(6)
    assert compile_graph(3, tri.nbr, tri.J, tri.S, tri.D).ucid == -1
    # one missing link (an open edge) breaks uniformity
    nbr = sq.nbr.copy()
    nbr[0, 0] = -1
    assert compile_graph(2, nbr, sq.J, sq.S, sq.D).ucid == -1
    # Ising: the folded coupling J*S_j must agree too
    S = np.ones(sq.n)
    S[3] = 2.0
    assert compile_graph(1, sq.nbr, sq.J[:, :, :1], np.ones(sq.n), np.zeros((sq.n, 3))).ucid >= 0
    assert compile_graph(1, sq.nbr, sq.J[:, :, :1], S, np.zeros((sq.n, 3))).ucid == -1
