"""Prepared runs (mcg_runner_*) against mcg_run and mcg_sweeps.

A prepared run plans, allocates and uploads once; each run then only queues
kernel launches and clears its own state with a fill kernel.  It must run the
same trajectory as mcg_run (which prepares, runs and releases every call):
spins, rungs and swap counts bit for bit, accumulators to FP64 summation
order.  Runs of one runner in chunks must match one long run.
"""
import ctypes as C

import numpy as np
import pytest

from mcsolver_amd import _lib as L
from mcsolver_amd import lattice as lat
from mcsolver_amd.engine import Batch, ReplicaSpec, compile_graph

from cases import _honeycomb

pytestmark = pytest.mark.gpu
SEED = 0x5EED2026


def _ladders():
    graphs, reps, betas, spins = [], [], [], []
    rng = np.random.default_rng(11)
    for gi, L_ in enumerate((16, 24)):
        Lt = lat.square(L_, J=-1.0, D=(0.05, 0.0, 0.0))
        graphs.append(compile_graph(2, Lt.nbr, Lt.J, Lt.S, Lt.D, pairs=Lt.corr_pairs))
        for r in range(8):
            reps.append(ReplicaSpec(gi, gi * 8, nrung=8, ladder=gi, rung=r))
            betas.append(1 / (0.7 + 0.07 * r))
            v = rng.normal(size=(Lt.n, 2))
            spins.append(v / np.linalg.norm(v, axis=1, keepdims=True))
    return Batch(2, graphs, reps, betas, [0.02] * len(betas), spins, nbins=4, nframes=40)


def _heis_stream():
    """Heisenberg honeycomb with 4 replicas on the streaming colour pass."""
    rng = np.random.default_rng(4)
    Lt = _honeycomb(12, 3)
    g = compile_graph(3, Lt.nbr, Lt.J, Lt.S, Lt.D, pairs=Lt.corr_pairs)
    spins = []
    for _ in range(4):
        v = rng.normal(size=(Lt.n, 3))
        spins.append(v / np.linalg.norm(v, axis=1, keepdims=True) * 1.5)
    return Batch(3, [g], [ReplicaSpec(0, k) for k in range(4)], [0.5, 0.9, 1.3, 2.0], [0.1] * 4, spins,
                 meas=L.MEAS_PAIRS, nbins=4, nframes=30)


def _wolff():
    rng = np.random.default_rng(6)
    Lt = lat.cubic(6, D=(0.0, 0.0, -0.05))
    g = compile_graph(3, Lt.nbr, Lt.J, Lt.S, Lt.D, pairs=Lt.corr_pairs)
    spins = []
    for _ in range(3):
        v = rng.normal(size=(Lt.n, 3))
        spins.append(v / np.linalg.norm(v, axis=1, keepdims=True))
    return Batch(3, [g], [ReplicaSpec(0, k) for k in range(3)], [0.6, 0.7, 0.8], [0.05] * 3, spins, nbins=4,
                 nframes=20, wolff=True)


def _snap(b):
    return (np.concatenate([b.replica_spins(k).ravel() for k in range(len(b.reps))]).astype(np.float32),
            b.accumulators(), b.bin_data(), b.pt_acc.cpu().numpy(), b.rung.cpu().numpy(), b.state_data(),
            b.error_flag())


def _same(a, b, rtol=1e-11):
    assert a[-1] == 0 and b[-1] == 0
    assert np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32)), "spins differ"
    assert np.array_equal(a[3], b[3]) and np.array_equal(a[4], b[4])
    np.testing.assert_allclose(b[1], a[1], rtol=rtol, atol=1e-9)
    np.testing.assert_allclose(b[2], a[2], rtol=rtol, atol=1e-9)


CASES = {
    # name: (builder, run kwargs)
    "ladders_persistent_split": (_ladders, dict(nthermal=10, nsweep=40, pt_every=1, kernel=L.KERNEL_LDS, parts=2)),
    "ladders_per_launch": (_ladders, dict(nthermal=10, nsweep=40, pt_every=2, kernel=L.KERNEL_LDS,
                                          no_persist=True)),
    "heisenberg_stream": (_heis_stream, dict(nthermal=5, nsweep=30, kernel=L.KERNEL_STREAM)),
    "wolff": (_wolff, dict(nthermal=4, nsweep=20, algorithm=L.WOLFF, clusters_per_sweep=4)),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_prepared_run_matches_mcg_run(gpu, case):
    mk, kw = CASES[case]
    print("case", case, flush=True)
    out = []
    for prepared in (False, True):
        b = mk()
        b.init()
        b.run(sweep0=0, prepared=prepared, **kw)
        out.append(_snap(b))
        b.close()
    _same(out[0], out[1])


@pytest.mark.parametrize("case", ["ladders_persistent_split", "heisenberg_stream", "wolff"])
def test_chunked_runs_match_one_run(gpu, case):
    """Three runs of one runner (sweep0 and frame0 carried on) equal one run
    of all the frames: every run starts its split counters, barrier words and
    published slots from zero again."""
    mk, kw = CASES[case]
    print("case", case, flush=True)
    kw = dict(kw)
    nth, nsw = kw.pop("nthermal"), kw.pop("nsweep")
    b = mk()
    b.init()
    b.run(nth, nsw, sweep0=0, **kw)
    one = _snap(b)
    b.close()
    c = mk()
    c.init()
    c.run(nth, 0, sweep0=0, **kw)
    done = 0
    for n in (7, 13, nsw - 20):
        c.run(0, n, sweep0=nth + done, frame0=done, **kw)
        done += n
    assert len(c._runners) <= 2  # one per plan shape
    _same(one, _snap(c), rtol=1e-9)
    c.close()


def test_sweeps_default_continues_the_stream(gpu):
    """Batch.run without sweep0 continues after the previous run, so two runs
    of n sweeps equal one run of 2n (the Philox counter is not replayed)."""
    b = _ladders()
    b.init()
    b.sweeps(0, 12, kernel=L.KERNEL_LDS)
    ref = _snap(b)[0]
    c = _ladders()
    c.init()
    c.sweeps(None, 6, kernel=L.KERNEL_LDS)
    c.sweeps(None, 6, kernel=L.KERNEL_LDS)
    assert np.array_equal(ref.view(np.uint32), _snap(c)[0].view(np.uint32))


def test_mcg_sweeps_matches_prepared_sweeps(gpu):
    b = _heis_stream()
    b.init()
    p = L.Plan()
    p.algorithm, p.kernel, p.sweeps_per_frame = L.METROPOLIS, L.KERNEL_STREAM, 1
    L.check(L.lib().mcg_sweeps(C.byref(b.b), C.byref(p), 3, 9, b._stream()), "mcg_sweeps")
    c = _heis_stream()
    c.init()
    c.sweeps(3, 9, kernel=L.KERNEL_STREAM)
    x, y = _snap(b), _snap(c)
    assert np.array_equal(x[0].view(np.uint32), y[0].view(np.uint32))
    np.testing.assert_allclose(y[5][:, :4], x[5][:, :4], rtol=1e-12, atol=1e-12)


def test_runner_refuses_another_plan_shape(gpu):
    b = _ladders()
    b.init()
    p = b.plan(0, 4, 1, pt_every=1, kernel=L.KERNEL_LDS, parts=2)
    h = b.runner(p)
    q = b.plan(0, 4, 1, pt_every=1, kernel=L.KERNEL_LDS, parts=1)
    with pytest.raises(RuntimeError, match="plan differs"):
        L.check(L.lib().mcg_runner_run(h, C.byref(b.b), C.byref(q), b._stream()), "mcg_runner_run")
    # a batch of another shape
    c = _heis_stream()
    with pytest.raises(RuntimeError, match="batch differs|plan differs"):
        L.check(L.lib().mcg_runner_run(h, C.byref(c.b), C.byref(p), c._stream()), "mcg_runner_run")
    L.check(L.lib().mcg_runner_run(h, C.byref(b.b), C.byref(p), b._stream()), "mcg_runner_run")
    assert b.error_flag() == 0
