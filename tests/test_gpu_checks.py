"""Failure detection and resume (forced from the host, every kernel exits
normally):

1. a spin corrupted through the host tensor between two runs: the next full
   FP64 recompute records |E_tracked - E_full| equal to the corruption's
   energy change in MCG_S_DRIFT and raises MCG_ERR_DRIFT;
2. NaN or Inf written into FP16 spin storage: the recompute raises
   MCG_ERR_NONFINITE (and the drop-in path raises a Python error);
3. spin lengths pushed off S: the recompute raises MCG_ERR_LENGTH;
4. a run checkpointed to a file, the batch discarded, a new batch restored
   from the file and continued: spins, ladder state and swap counts are bit
   for bit those of the uninterrupted run, accumulators equal to FP64
   summation order;
5. a clean production-length run raises nothing and its drift stays small.
"""
import os

import numpy as np
import pytest
import torch

from mcsolver_amd import _lib as L
from mcsolver_amd import lattice as lat
from mcsolver_amd.engine import Batch, ReplicaSpec, compile_graph
from oracle import mcref

from cases import _honeycomb

pytestmark = pytest.mark.gpu


def _heis(store="f32", R=2, recompute_every=10, L_=6, kernel_lds=True):
    Lt = _honeycomb(L_, 3, S=1.0)
    g = compile_graph(3, Lt.nbr, Lt.J, Lt.S, Lt.D)
    rng = np.random.default_rng(8)
    spins = []
    for _ in range(R):
        v = rng.normal(size=(Lt.n, 3))
        spins.append(v / np.linalg.norm(v, axis=1, keepdims=True))
    b = Batch(3, [g], [ReplicaSpec(0, k) for k in range(R)], [1.2] * R, [0.1] * R, spins, store=store,
              recompute_every=recompute_every, nbins=1, nframes=1)
    b.init()
    return b, g, Lt


def _pool_index(b, g, k, site, comp):
    """Element of the spin pool holding component comp of site of replica k
    (replicas store their spins in colour order)."""
    p = int(np.nonzero(g.col_site == site)[0][0])
    return b.spin_off[k] + comp * g.n + p


def test_corrupted_spin_is_flagged_by_the_recompute(gpu):
    b, g, Lt = _heis(recompute_every=10)
    b.sweeps(0, 20)
    assert b.error_flag() == 0
    s = b.replica_spins(0).astype(np.float32)
    e0 = mcref.cb_energy(3, Lt.nbr, Lt.J, Lt.S, Lt.D, 0.1, np.ascontiguousarray(s.T))[0]
    site = 5
    new = -s[site]                     # still a unit vector: only the energy is wrong
    for c in range(3):
        b.spins[_pool_index(b, g, 0, site, c)] = float(new[c])
    s2 = s.copy()
    s2[site] = new
    e1 = mcref.cb_energy(3, Lt.nbr, Lt.J, Lt.S, Lt.D, 0.1, np.ascontiguousarray(s2.T))[0]
    dE = e1 - e0
    print("corruption changes E by %.4f" % dE)
    assert abs(dE) > 0.5
    b.sweeps(None, 10)
    st = b.state_data()
    flag = b.error_flag()
    print("drift replica 0: %.6f, replica 1: %.3g, flag 0x%x" % (st[0, L.S_DRIFT], st[1, L.S_DRIFT], flag))
    assert abs(st[0, L.S_DRIFT] - abs(dE)) < 1e-3 * abs(dE) + 1e-4
    assert st[1, L.S_DRIFT] < 1e-3
    assert flag & L.ERR_DRIFT and not flag & (L.ERR_NONFINITE | L.ERR_LENGTH)


@pytest.mark.parametrize("bad", [float("nan"), float("inf")])
def test_nonfinite_fp16_spin_trips_the_guard(gpu, bad):
    b, g, Lt = _heis(store="f16", recompute_every=5)
    b.sweeps(0, 5)
    assert b.error_flag() == 0
    b.spins[_pool_index(b, g, 1, 7, 2)] = bad
    b.sweeps(None, 5)
    flag = b.error_flag()
    print("flag 0x%x" % flag)
    assert flag & L.ERR_NONFINITE


def test_spin_length_drift_trips_the_check(gpu):
    b, g, Lt = _heis(recompute_every=1)
    b.sweeps(0, 3)
    assert b.error_flag() == 0
    o = b.spin_off[0]
    b.spins[o:o + 3 * g.n] *= 1.02      # every spin of replica 0 2 % too long
    b.sweeps(None, 1)
    flag = b.error_flag()
    print("flag 0x%x" % flag)
    assert flag & L.ERR_LENGTH


def test_dropin_raises_on_device_errors(gpu):
    from mcsolver_amd.dropin import check_health
    b, g, Lt = _heis(store="f16", recompute_every=2)
    b.spins[_pool_index(b, g, 0, 3, 0)] = float("nan")
    b.sweeps(0, 2)
    with pytest.raises(RuntimeError, match="non-finite"):
        check_health(b)


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
    return Batch(2, graphs, reps, betas, [0.02] * len(betas), spins, nbins=4, nframes=40, recompute_every=7)


def _snap(b):
    return (np.concatenate([b.replica_spins(k).ravel() for k in range(len(b.reps))]).astype(np.float32),
            b.accumulators(), b.bin_data(), b.pt_acc.cpu().numpy(), b.rung.cpu().numpy(), b.who.cpu().numpy(),
            b.state_data(), b.error_flag())


@pytest.mark.parametrize("pt_every", [1, 3])
@pytest.mark.parametrize("kernel", [L.KERNEL_LDS, L.KERNEL_STREAM])
def test_checkpoint_resume_is_bit_identical(gpu, kernel, pt_every, tmp_path):
    """Tempering ladders, frames every sweep, FP64 recompute every 7 sweeps:
    10 thermal + 40 measured sweeps in one run, against 10 + 15, checkpoint
    to a file, a new batch restored from it, then 25 more.  Recomputes and
    swaps follow the global sweep index, so the split at sweep 25 (not a
    multiple of 7 or 3) changes nothing."""
    a = _ladders()
    a.init()
    a.run(10, 40, 1, pt_every=pt_every, kernel=kernel, sweep0=0)
    ref = _snap(a)
    a.close()
    del a
    b = _ladders()
    b.init()
    b.run(10, 15, 1, pt_every=pt_every, kernel=kernel, sweep0=0)
    f = os.path.join(str(tmp_path), "ckpt.npz")
    np.savez(f, **b.checkpoint())
    b.close()
    del b
    torch.cuda.synchronize()
    c = _ladders()
    c.restore(dict(np.load(f)))
    assert c._next_sweep == 25
    c.run(0, 25, 1, pt_every=pt_every, kernel=kernel, frame0=15)
    got = _snap(c)
    assert ref[-1] == 0 and got[-1] == 0
    assert np.array_equal(ref[0].view(np.uint32), got[0].view(np.uint32))
    for x, y in zip(ref[3:6], got[3:6]):
        assert np.array_equal(x, y)
    np.testing.assert_allclose(got[1], ref[1], rtol=1e-11, atol=1e-9)
    np.testing.assert_allclose(got[2], ref[2], rtol=1e-11, atol=1e-9)


def test_clean_run_raises_nothing(gpu):
    """FP32 and FP16 storage, 2000 sweeps with the production recompute
    period: no error bit, drift far below its tolerance."""
    for store in ("f32", "f16"):
        b, g, Lt = _heis(store=store, R=4, recompute_every=L.RECOMPUTE_EVERY, L_=16)
        b.sweeps(0, 2000)
        st = b.state_data()
        print(store, "max drift %.3g (n = %d)" % (st[:, L.S_DRIFT].max(), g.n))
        assert b.error_flag() == 0
        assert st[:, L.S_DRIFT].max() < (1e-5 if store == "f32" else 1e-3) * g.n
