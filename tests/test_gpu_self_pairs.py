"""Self pairs: on one-orbital lattices every correlation pair is (i, i)
(mcMain.py's default ki_s = ki_t, overLat 0), and the frame takes the pair
sums from the tracked magnetisation (sum s_i) and the spin lengths
(sum s_i.s_i = sum S_i^2) instead of reading the spins.  One measured frame
on each kernel path must equal the sums over the spins it leaves, to FP32
tracking accuracy."""
import numpy as np
import pytest

from mcsolver_amd import _lib as L
from mcsolver_amd import lattice as lat
from mcsolver_amd.engine import Batch, ReplicaSpec, compile_graph

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kernel", [L.KERNEL_LDS, L.KERNEL_STREAM, L.KERNEL_PLAIN])
@pytest.mark.parametrize("nc", [2, 3])
def test_self_pair_sums_equal_spin_sums(gpu, nc, kernel):
    Lt = lat.cubic(8, J=-1.0, S=1.5, D=(0.0, 0.0, -0.1)) if nc == 3 else lat.square(24, J=-1.0, S=1.5)
    g = compile_graph(nc, Lt.nbr, Lt.J, Lt.S, Lt.D, pairs=Lt.corr_pairs)
    rng = np.random.default_rng(4)
    betas = [0.5, 1.2, 3.0]
    spins = []
    for _ in betas:
        v = rng.normal(size=(Lt.n, nc))
        spins.append(1.5 * v / np.linalg.norm(v, axis=1, keepdims=True))
    b = Batch(nc, [g], [ReplicaSpec(0, k) for k in range(len(betas))], betas, [0.05] * 3, spins,
              meas=L.MEAS_PAIRS, nbins=1, nframes=1)
    assert g.struct.self_pairs == 1 and abs(g.struct.pair_s2 - Lt.n * 2.25) < 1e-9
    b.init()
    b.run(7, 1, 1, kernel=kernel, sweep0=0)
    assert b.error_flag() == 0
    A = b.accumulators()
    for k in range(len(betas)):
        s = b.replica_spins(k).astype(np.float64)          # [n, nc]
        tot = s.sum(axis=0)
        np.testing.assert_allclose(A[k, L.A_SI:L.A_SI + nc], np.abs(tot) / Lt.n, rtol=1e-5, atol=1e-6)
        np.testing.assert_allclose(A[k, L.A_SJ:L.A_SJ + nc], np.abs(tot) / Lt.n, rtol=1e-5, atol=1e-6)
        np.testing.assert_allclose(A[k, L.A_SIJ], (s * s).sum() / Lt.n, rtol=1e-6)
