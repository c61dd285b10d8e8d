"""mcg_run_noalloc (no device allocation; scratch in a per-device file-scope
array) against mcg_run, through the C ABI.

Two calls in a row on one device, the second with a smaller batch, each
checked against mcg_run on an identical fresh batch: spins, rungs and swap
counts bit for bit, accumulators to FP64 summation order.  The second call
finds the first call's contents in the array, so this also checks that the
run state is cleared.  A batch whose tables do not fit the array is refused
before anything is launched.
"""
import ctypes as C

import numpy as np
import pytest

from mcsolver_amd import _lib as L
from mcsolver_amd import lattice as lat
from mcsolver_amd.engine import Batch, ReplicaSpec, compile_graph

pytestmark = pytest.mark.gpu


def _ladders(sizes, rungs, seed):
    graphs, reps, betas, spins = [], [], [], []
    rng = np.random.default_rng(seed)
    for gi, L_ in enumerate(sizes):
        Lt = lat.square(L_, J=-1.0, D=(0.05, 0.0, 0.0))
        graphs.append(compile_graph(2, Lt.nbr, Lt.J, Lt.S, Lt.D, pairs=Lt.corr_pairs))
        for r in range(rungs):
            reps.append(ReplicaSpec(gi, gi * rungs, nrung=rungs, ladder=gi, rung=r))
            betas.append(1 / (0.7 + 0.5 * r / (rungs - 1)))
            v = rng.normal(size=(Lt.n, 2))
            spins.append(v / np.linalg.norm(v, axis=1, keepdims=True))
    return Batch(2, graphs, reps, betas, [0.02] * len(betas), spins, nbins=4, nframes=30)


def _heis(n_rep, L_, seed):
    rng = np.random.default_rng(seed)
    Lt = lat.cubic(L_, D=(0.0, 0.0, -0.05))
    g = compile_graph(3, Lt.nbr, Lt.J, Lt.S, Lt.D, pairs=Lt.corr_pairs)
    spins = []
    for _ in range(n_rep):
        v = rng.normal(size=(Lt.n, 3))
        spins.append(v / np.linalg.norm(v, axis=1, keepdims=True))
    return Batch(3, [g], [ReplicaSpec(0, k) for k in range(n_rep)], list(np.linspace(0.5, 1.5, n_rep)),
                 [0.1] * n_rep, spins, meas=L.MEAS_PAIRS, nbins=4, nframes=30)


def _snap(b):
    return (np.concatenate([b.replica_spins(k).ravel() for k in range(len(b.reps))]).astype(np.float32),
            b.accumulators(), b.bin_data(), b.pt_acc.cpu().numpy(), b.rung.cpu().numpy(), b.error_flag())


def _run(b, p, noalloc):
    f = L.lib().mcg_run_noalloc if noalloc else L.lib().mcg_run
    L.check(f(C.byref(b.b), C.byref(p), b._stream()), "mcg_run_noalloc" if noalloc else "mcg_run")


# (builder, plan kwargs): each pair is a larger call, then a smaller one
PAIRS = {
    "ladders_persistent_split": [
        (lambda: _ladders((16, 24, 32), 8, 1), dict(nthermal=8, nsweep=30, pt_every=1, kernel=L.KERNEL_LDS, parts=2)),
        (lambda: _ladders((16,), 4, 2), dict(nthermal=5, nsweep=20, pt_every=1, kernel=L.KERNEL_LDS, parts=2)),
    ],
    "stream_then_lds": [
        (lambda: _heis(6, 8, 3), dict(nthermal=4, nsweep=30, kernel=L.KERNEL_STREAM)),
        (lambda: _heis(2, 6, 4), dict(nthermal=3, nsweep=10, kernel=L.KERNEL_AUTO)),
    ],
}


@pytest.mark.parametrize("case", sorted(PAIRS))
def test_noalloc_matches_mcg_run_on_two_calls(gpu, case):
    for step, (mk, kw) in enumerate(PAIRS[case]):
        print("case", case, "call", step, flush=True)
        out = []
        for noalloc in (False, True):
            b = mk()
            b.init()
            _run(b, b.plan(sweep0=0, **kw), noalloc)
            out.append(_snap(b))
        a, c = out
        assert a[-1] == 0 and c[-1] == 0
        assert np.array_equal(a[0].view(np.uint32), c[0].view(np.uint32)), "spins differ"
        assert np.array_equal(a[3], c[3]) and np.array_equal(a[4], c[4])
        np.testing.assert_allclose(c[1], a[1], rtol=1e-11, atol=1e-9)
        np.testing.assert_allclose(c[2], a[2], rtol=1e-11, atol=1e-9)


def test_noalloc_refuses_timing_and_oversized_runs(gpu):
    b = _heis(2, 4, 5)
    b.init()
    t = (C.c_double * 4)()
    with pytest.raises(RuntimeError, match="timing"):
        _run(b, b.plan(0, 2, timing=t), True)
    # 160,000 replicas of a 2x2 Ising lattice: the work-item tables alone
    # (24 B per replica) exceed the array
    Lt = lat.square(2, ising=True)
    g = compile_graph(1, Lt.nbr, Lt.J[:, :, :1], Lt.S, Lt.D)
    n_rep = 160_000
    big = Batch(1, [g], [ReplicaSpec(0, k) for k in range(n_rep)], [0.5] * n_rep, [0.0] * n_rep,
                [np.ones(Lt.n)] * n_rep, nbins=1)
    with pytest.raises(RuntimeError, match="bytes of device scratch"):
        _run(big, big.plan(0, 1), True)
