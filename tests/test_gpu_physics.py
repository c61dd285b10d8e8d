"""Engine observables against closed forms (tests/exact.py).

Error bars: independent replicas per temperature (spread / sqrt(R)); the
acceptance is 4 sigma plus a small absolute floor for the FP32 arithmetic.
"""
import numpy as np
import pytest

import exact
from mcsolver_amd import _lib as L
from mcsolver_amd import lattice as lat
from mcsolver_amd.engine import Batch, ReplicaSpec, compile_graph

pytestmark = pytest.mark.gpu


def _run(nc, Lt, temps, R, nth, nsw, h=0.0, store="f32", meas=0, kernel=L.KERNEL_AUTO, init="x", seed=7):
    J = Lt.J if nc > 1 else Lt.J[:, :, :1]
    g = compile_graph(nc, Lt.nbr, J, Lt.S, Lt.D, pairs=Lt.corr_pairs)
    reps, beta, hs, inits = [], [], [], []
    rng = np.random.default_rng(seed)
    for T in temps:
        for _ in range(R):
            reps.append(ReplicaSpec(0, len(reps)))
            beta.append(1.0 / T)
            hs.append(h)
            if nc == 1:
                inits.append(np.ones(Lt.n))
            elif init == "x":
                v = np.zeros((Lt.n, nc)); v[:, 0] = np.abs(Lt.S); inits.append(v)
            else:
                v = rng.normal(size=(Lt.n, nc)); inits.append(v / np.linalg.norm(v, axis=1, keepdims=True))
    b = Batch(nc, [g], reps, beta, hs, inits, store=store, meas=meas, nbins=8, nframes=nsw, seed=seed)
    b.init()
    b.run(nth, nsw, 1, kernel=kernel)
    A = b.accumulators()
    assert b.error_flag() == 0
    n = A[:, L.A_COUNT]
    e = (A[:, L.A_E] / n).reshape(len(temps), R)
    e2 = (A[:, L.A_E2] / n).reshape(len(temps), R)
    return b, e, e2, A


@pytest.mark.parametrize("kernel", [L.KERNEL_LDS, L.KERNEL_STREAM])
def test_ising_torus_kaufman(gpu, kernel):
    Ls = 16
    temps = [2.0, 2.269, 2.6]
    Lt = lat.square(Ls, J=-1.0, ising=True)
    b, e, e2, _ = _run(1, Lt, temps, 8, 2000, 20000, kernel=kernel)
    N = Lt.n
    for k, T in enumerate(temps):
        e_ex, c_ex = exact.ising_e_c(Ls, T)
        c = N * (e2[k] - e[k] ** 2) / T ** 2
        print("T=%.3f e=%.5f+-%.5f exact %.5f   c=%.4f+-%.4f exact %.4f" % (
            T, e[k].mean(), e[k].std(ddof=1) / np.sqrt(8), e_ex, c.mean(), c.std(ddof=1) / np.sqrt(8), c_ex))
        assert abs(e[k].mean() - e_ex) < 4 * e[k].std(ddof=1) / np.sqrt(8) + 1e-4
        assert abs(c.mean() - c_ex) < 4 * c.std(ddof=1) / np.sqrt(8) + 1e-3


@pytest.mark.parametrize("model,nc,T", [("xy", 2, 0.8), ("heis", 3, 0.7), ("xy", 2, 2.0), ("heis", 3, 1.5)])
def test_ring_energy(gpu, model, nc, T):
    Lt = lat.build_lattice(128, 1, 1, np.eye(3), [[0, 0, 0]], [lat.Bond(0, 0, (1, 0, 0), [-1.0] * 3)], [1.0])
    _, e, _, _ = _run(nc, Lt, [T], 16, 500, 4000)
    want = -exact.ring_bond_energy(model, 1.0 / T)
    assert abs(e[0].mean() - want) < 4 * e[0].std(ddof=1) / 4 + 2e-4, (e[0].mean(), want)


@pytest.mark.parametrize("nc,store", [(3, "f32"), (3, "f16"), (2, "f32")])
def test_isolated_spins_anisotropy_field(gpu, nc, store):
    """No bonds: E per site and the field-axis moment vs quadrature."""
    D, h, T = (0.1, -0.2, -0.5), 0.3, 0.8
    if nc == 2:
        D = (-0.3, 0.2, 0.0)
    Lt = lat.build_lattice(32, 32, 1, np.eye(3), [[0, 0, 0]], [], [1.0], [list(D)])
    b, e, _, A = _run(nc, Lt, [T], 8, 200, 2000, h=h, store=store)
    e_ex, m_ex = exact.single_spin(D, h, 1 / T, nc=nc)
    m = A[:, L.A_HT] / A[:, L.A_COUNT]
    tol = 2e-3 if store == "f16" else 2e-4
    assert abs(e[0].mean() - e_ex) < 4 * e[0].std(ddof=1) / np.sqrt(8) + tol, (e[0].mean(), e_ex)
    assert abs(m.mean() - m_ex) < 4 * m.std(ddof=1) / np.sqrt(8) + tol, (m.mean(), m_ex)


def test_tensor_orientation_nonreciprocal(gpu):
    """Initial FP64 energy of a random non-reciprocal graph equals the
    reference formula E = 1/2 sum_i sum_l diagonalDot(s_i, s_j, J_il) +
    sum_a D_a s_a^2 - h s_z (heisenbergLib.c:77-87, 238-253): pins the order
    xx,yy,zz,xy,xz,yz,yx,zx,zy and J (not J^T) on the source side."""
    rng = np.random.default_rng(4)
    n, z = 96, 5
    nbr = rng.integers(0, n, (n, z))
    J = rng.normal(size=(n, z, 9))
    S = np.ones(n)
    D = rng.normal(size=(n, 3)) * 0.3
    h = 0.37
    s = rng.normal(size=(n, 3))
    s /= np.linalg.norm(s, axis=1, keepdims=True)
    g = compile_graph(3, nbr, J, S, D)
    b = Batch(3, [g], [ReplicaSpec(0, 0)], [1.0], [h], [s])
    b.init()
    E = b.state_data()[0, L.S_E]
    sf = s.astype(np.float32).astype(np.float64)
    ref = 0.0
    for i in range(n):
        for l in range(z):
            a, c, jj = sf[i], sf[nbr[i, l]], J[i, l]
            ref += 0.5 * (a[0] * c[0] * jj[0] + a[1] * c[1] * jj[1] + a[2] * c[2] * jj[2] + a[0] * c[1] * jj[3] +
                          a[0] * c[2] * jj[4] + a[1] * c[2] * jj[5] + a[1] * c[0] * jj[6] + a[2] * c[0] * jj[7] +
                          a[2] * c[1] * jj[8])
        ref += D[i, 0] * sf[i, 0] ** 2 + D[i, 1] * sf[i, 1] ** 2 + D[i, 2] * sf[i, 2] ** 2 - h * sf[i, 2]
    assert abs(E - ref) < 1e-4 * n
    # and the transposed reading would differ
    Jt = J[:, :, [0, 1, 2, 6, 7, 8, 3, 4, 5]]
    alt = sum(0.5 * sf[i] @ np.array([[Jt[i, l, 0], Jt[i, l, 3], Jt[i, l, 4]], [Jt[i, l, 6], Jt[i, l, 1], Jt[i, l, 5]],
                                       [Jt[i, l, 7], Jt[i, l, 8], Jt[i, l, 2]]]) @ sf[nbr[i, l]]
              for i in range(n) for l in range(z))
    alt += sum(D[i] @ sf[i] ** 2 - h * sf[i, 2] for i in range(n))
    assert abs(alt - ref) > 1.0
