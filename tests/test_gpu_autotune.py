"""The measured plan choice (mcsolver_amd/autotune.py): it leaves the batch as
it found it, and the run with the plan it picks is the run with the default
plan (same spins and swaps; frame observables to FP32 rounding)."""
import numpy as np
import pytest

from mcsolver_amd import lattice as lat
from mcsolver_amd.engine import Batch, ReplicaSpec, compile_graph

pytestmark = pytest.mark.gpu


def _batch(seed=21):
    """XY ladders on L = 16 (8 rungs) and L = 96 (4 rungs, large enough to be
    split automatically), with pair correlations measured."""
    import torch
    torch.cuda.set_device(0)
    rng = np.random.default_rng(seed)
    graphs, reps, betas, spins = [], [], [], []
    for gi, (L_, nr) in enumerate(((16, 8), (96, 4))):
        Lt = lat.square(L_, J=-1.0)
        graphs.append(compile_graph(2, Lt.nbr, Lt.J, Lt.S, Lt.D))
        row0 = len(betas)
        for r in range(nr):
            reps.append(ReplicaSpec(gi, row0, nrung=nr, ladder=gi, rung=r))
            betas.append(1 / (0.8 + 0.05 * r))
            v = rng.normal(size=(Lt.n, 2))
            spins.append(v / np.linalg.norm(v, axis=1, keepdims=True))
    b = Batch(2, graphs, reps, betas, [0.01] * len(betas), spins, nbins=4, nframes=30)
    b.init()
    return b


def _snap(b):
    return [x.cpu().numpy().copy() for x in (b.spins, b.state, b.acc, b.bins, b.pt_acc, b.rung, b.who, b.err)]


def test_autotune_restores_batch_and_keeps_trajectory(gpu):
    from mcsolver_amd.autotune import autotune
    b = _batch()
    b.run(5, 0, 1, pt_every=1)  # some history, so that restoring is not restoring zeros
    before = _snap(b)
    r = autotune(b, 1, 1, trial_sweeps=12)
    after = _snap(b)
    for x, y in zip(before, after):
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8))
    ok = [t for t in r.trials if t.ms_per_sweep is not None]
    assert len(ok) >= 2 and r.best in [t.options for t in ok]
    print("autotune trials:", [(t.options, t.ms_per_sweep) for t in r.trials], flush=True)
    # production run with the chosen plan vs the default plan on a fresh batch
    b.run(0, 30, 1, pt_every=1, sweep0=5, **r.best)
    d = _batch()
    d.run(5, 0, 1, pt_every=1)
    d.run(0, 30, 1, pt_every=1, sweep0=5)
    assert b.error_flag() == 0 and d.error_flag() == 0
    assert np.array_equal(b.spins.cpu().numpy().view(np.uint32), d.spins.cpu().numpy().view(np.uint32))
    for x, y in ((b.rung, d.rung), (b.who, d.who), (b.pt_acc, d.pt_acc)):
        assert np.array_equal(x.cpu().numpy(), y.cpu().numpy())
    # the tracked E and M are FP64 sums of per-thread FP32 partial sums whose
    # grouping depends on how sites are dealt to threads: FP32 rounding level
    for x, y in ((b.accumulators(), d.accumulators()), (b.bin_data(), d.bin_data())):
        scale = np.abs(y).reshape(-1, y.shape[-1]).max(axis=0)  # per observable column
        assert (np.abs(x - y) <= 2e-6 * scale).all(), np.abs(x - y).max()
