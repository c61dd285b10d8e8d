"""GPU sweep == CPU checker, bit for bit.

Every kernel path (LDS-resident persistent, streaming colour pass, plain
one-launch-per-replica) runs the same colour-ordered Metropolis trajectory
as oracle/checker.c: same Philox counters, same FP32 operation sequence.
Spins must be identical; tracked energies agree to FP64 summation order
(exact for Ising with integer couplings).
"""
import numpy as np
import pytest

from mcsolver_amd import _lib as L
from mcsolver_amd import lattice as lat
from mcsolver_amd.engine import Batch, ReplicaSpec, compile_graph
from oracle import mcref

from cases import _honeycomb, _triangular_dmi

pytestmark = pytest.mark.gpu
SEED = 0x5EED2026


def _lattice_tables(L_, nc):
    J = L_.J if nc > 1 else L_.J[:, :, :1]
    return L_.nbr, J, L_.S, L_.D


def _mk_batch(nc, Lt, betas, h, store="f32", init=None, wolff=False, seed=SEED, **kw):
    nbr, J, S, D = _lattice_tables(Lt, nc)
    g = compile_graph(nc, nbr, J, S, D, pairs=Lt.corr_pairs)
    n = Lt.n
    rng = np.random.default_rng(5)
    reps, spins = [], []
    for k in range(len(betas)):
        reps.append(ReplicaSpec(0, k))
        if init is not None:
            spins.append(init[k])
        elif nc == 1:
            spins.append(np.where(rng.random(n) < 0.5, 1.0, -1.0))
        else:
            v = rng.normal(size=(n, nc))
            v /= np.linalg.norm(v, axis=1, keepdims=True)
            v *= np.abs(S)[:, None]
            if store == "f16":
                v = v.astype(np.float16).astype(np.float64)
            spins.append(v)
    b = Batch(nc, [g], reps, betas, [h] * len(betas), spins, store=store, seed=seed, wolff=wolff, **kw)
    b.init()
    return b, g, spins


def _cpu_spins(nc, spins, store):
    out = np.zeros((len(spins), nc, spins[0].shape[0]), np.float32)
    for k, s in enumerate(spins):
        a = np.asarray(s, np.float64).reshape(len(s), -1)
        if nc == 1:
            out[k, 0] = np.where(a[:, 0] >= 0, 1.0, -1.0)
        else:
            out[k] = a[:, :nc].T.astype(np.float32)
    return out


def _gpu_spins(b, nrep):
    return np.stack([b.replica_spins(k).T.astype(np.float32) for k in range(nrep)])


CASES = {
    "ising_sq16": (1, lambda: lat.square(16, ising=True), [0.2, 0.44, 0.6, 1.0], 0.05),
    "ising_2orb": (1, lambda: __import__("cases")._ising_2orb(8), [0.3, 0.5], 0.2),
    "xy_sq16": (2, lambda: lat.square(16, J=-1.0, D=(0.1, -0.05, 0.0)), [0.5, 1.1, 2.0, 3.0], 0.1),
    "xy_sq12_noD": (2, lambda: lat.square(12, J=-1.0), [0.7, 1.4], 0.0),
    # 144 groups per colour: each replica's partial last round goes to the pool
    "xy_sq24_pool": (2, lambda: lat.square(24, J=-1.0), [0.6, 0.9, 1.2, 2.0], 0.05),
    "heis_cubic4_noD": (3, lambda: lat.cubic(4, J=-1.0), [0.6, 1.6], 0.05),
    "heis_honeycomb": (3, lambda: _honeycomb(8, 3), [0.7, 1.5], 0.2),
    "heis_tri_dmi": (3, lambda: _triangular_dmi(12), [1.0, 4.0], 0.3),
}


@pytest.mark.parametrize("kernel", [L.KERNEL_LDS, L.KERNEL_STREAM, L.KERNEL_PLAIN])
@pytest.mark.parametrize("case", sorted(CASES))
def test_metropolis_bitexact(gpu, case, kernel):
    nc, mk, betas, h = CASES[case]
    Lt = mk()
    print("case", case, "kernel", kernel, "n", Lt.n, flush=True)
    b, g, init = _mk_batch(nc, Lt, betas, h)
    nsw = 12
    b.sweeps(0, nsw, kernel=kernel)
    got = _gpu_spins(b, len(betas))
    ref = _cpu_spins(nc, init, "f32")
    nbr, J, S, D = _lattice_tables(Lt, nc)
    dE, dM = mcref.cb_sweeps(nc, 0 if nc > 1 else 2, nbr, J, S, D, g.colour, betas, [h] * len(betas),
                             np.arange(len(betas), dtype=np.uint32), SEED, 0, nsw, ref)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), "spins differ"
    st = b.state_data()
    for k in range(len(betas)):
        e_cpu = mcref.cb_energy(nc, nbr, J, S, D, h, ref[k])
        # tracked energy (initial FP64 recompute + accepted dE) vs a fresh recompute
        assert abs(st[k, L.S_E] - e_cpu[0]) <= 1e-3 * max(1.0, abs(e_cpu[0])), (k, st[k, L.S_E], e_cpu[0])
        np.testing.assert_allclose(st[k, L.S_M:L.S_M + 3], e_cpu[1:], rtol=1e-4, atol=1e-3)
    assert b.error_flag() == 0


def test_lds_multi_replica_packing(gpu):
    """Many small replicas packed into one LDS block follow their own streams."""
    nc, Lt = 2, lat.square(8, J=-1.0)
    betas = list(np.linspace(0.3, 2.5, 16))
    b, g, init = _mk_batch(nc, Lt, betas, 0.0)
    b.sweeps(0, 7, kernel=L.KERNEL_LDS, block=1024)
    got = _gpu_spins(b, len(betas))
    ref = _cpu_spins(nc, init, "f32")
    nbr, J, S, D = _lattice_tables(Lt, nc)
    mcref.cb_sweeps(nc, 0, nbr, J, S, D, g.colour, betas, [0.0] * 16, np.arange(16, dtype=np.uint32), SEED, 0, 7, ref)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_fp16_storage_bitexact(gpu):
    nc, Lt = 3, _honeycomb(8, 4)
    betas = [0.8, 2.0]
    b, g, init = _mk_batch(nc, Lt, betas, 0.1, store="f16")
    b.sweeps(0, 9, kernel=L.KERNEL_LDS)
    got = _gpu_spins(b, 2)
    ref = _cpu_spins(nc, init, "f16")
    nbr, J, S, D = _lattice_tables(Lt, nc)
    mcref.cb_sweeps(nc, 1, nbr, J, S, D, g.colour, betas, [0.1, 0.1], np.arange(2, dtype=np.uint32), SEED, 0, 9, ref)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("case", ["ising_sq16", "xy_sq16", "heis_honeycomb"])
def test_wolff_bitexact(gpu, case):
    nc, mk, betas, h = CASES[case]
    Lt = mk()
    betas = betas[:2]
    print("wolff case", case, flush=True)
    b, g, init = _mk_batch(nc, Lt, betas, h, wolff=True)
    ncl = 6
    b.sweeps(0, 2, algorithm=L.WOLFF, clusters_per_sweep=ncl)
    got = _gpu_spins(b, len(betas))
    ref = _cpu_spins(nc, init, "f32")
    nbr, J, S, D = _lattice_tables(Lt, nc)
    for k in range(len(betas)):
        one = np.ascontiguousarray(ref[k])
        mcref.cb_wolff(nc, 0 if nc > 1 else 2, nbr, J, S, D, g.colour, betas[k], h, k, SEED, 0, 2 * ncl, one)
        ref[k] = one
    if nc == 1:
        assert np.array_equal(got, ref)
    else:
        # the cluster acceptance sums FP32 terms in FP64 in a different order;
        # a borderline accept could flip, so allow none but report clearly
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


WOLFF_ZERO = {
    # name: (nc, lattice, betas, h, correction is identically zero)
    "heis_cubic4_noD_h0": (3, lambda: lat.cubic(4, J=-1.0), [0.6, 1.6], 0.0, True),
    "xy_sq12_noD_h0": (2, lambda: lat.square(12, J=-1.0), [0.7, 1.4], 0.0, True),
    "ising_sq16_h0": (1, lambda: lat.square(16, ising=True), [0.3, 0.6], 0.0, True),
    # isotropic exchange but a single-ion term, or a field: evaluated
    "heis_cubic4_D_h0": (3, lambda: lat.cubic(4, J=-1.0, D=(0.0, 0.0, -0.4)), [0.6, 1.6], 0.0, False),
    "heis_cubic4_noD_h": (3, lambda: lat.cubic(4, J=-1.0), [0.6, 1.6], 0.3, False),
    "xy_sq12_D_h0": (2, lambda: lat.square(12, J=-1.0, D=(0.3, 0.0, 0.0)), [0.7, 1.4], 0.0, False),
}


@pytest.mark.parametrize("case", sorted(WOLFF_ZERO))
def test_wolff_zero_correction_rule_bitexact(gpu, case):
    """The Wolff cluster acceptance skips the correction energy exactly when
    it is identically zero (isotropic exchange, no single-ion term, h = 0;
    heisenbergLib.c:403-423): engine and checker agree on that rule (spins
    bit for bit), every cluster is then accepted (S_ACC == S_TRY), and with
    a single-ion term or a field the correction is evaluated by both."""
    nc, mk, betas, h, zero = WOLFF_ZERO[case]
    Lt = mk()
    print("wolff case", case, flush=True)
    b, g, init = _mk_batch(nc, Lt, betas, h, wolff=True)
    ncl = 8
    b.sweeps(0, 3, algorithm=L.WOLFF, clusters_per_sweep=ncl)
    got = _gpu_spins(b, len(betas))
    ref = _cpu_spins(nc, init, "f32")
    nbr, J, S, D = _lattice_tables(Lt, nc)
    for k in range(len(betas)):
        one = np.ascontiguousarray(ref[k])
        mcref.cb_wolff(nc, 0 if nc > 1 else 2, nbr, J, S, D, g.colour, betas[k], h, k, SEED, 0, 3 * ncl, one)
        ref[k] = one
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    st = b.state_data()
    print("accepted / grown spins:", st[:, L.S_ACC], st[:, L.S_TRY])
    assert (st[:, L.S_TRY] > 0).all()
    if zero:
        assert np.array_equal(st[:, L.S_ACC], st[:, L.S_TRY])
    else:
        assert (st[:, L.S_ACC] <= st[:, L.S_TRY]).all()
    assert b.error_flag() == 0


@pytest.mark.parametrize("no_persist", [False, True])
def test_parallel_tempering_bitexact(gpu, no_persist):
    """Ising ladder, swap every sweep: permutation, rungs and spins match
    (persistent launch with tempering inside, and one launch per sweep)."""
    nc, Lt = 1, lat.square(8, ising=True)
    nT = 8
    betas = list(np.linspace(0.25, 0.65, nT).astype(np.float32))
    nbr, J, S, D = _lattice_tables(Lt, nc)
    g = compile_graph(1, nbr, J, S, D, pairs=Lt.corr_pairs)
    rng = np.random.default_rng(9)
    init = [np.where(rng.random(Lt.n) < 0.5, 1.0, -1.0) for _ in range(nT)]
    reps = [ReplicaSpec(0, 0, nrung=nT, ladder=0, rung=k) for k in range(nT)]
    b = Batch(1, [g], reps, betas, [0.0] * nT, init, seed=SEED)
    b.init()
    assert b.ising_table() is not None  # J = 1, h = 0: the exact acceptance
    nsw = 10
    b.run(0, nsw, 1, pt_every=1, kernel=L.KERNEL_LDS, no_persist=no_persist)
    got = _gpu_spins(b, nT)
    # CPU: per sweep, every replica at its current rung's beta, then swap
    ref = _cpu_spins(1, init, "f32")
    who = np.arange(nT, dtype=np.int32)
    rung = np.arange(nT, dtype=np.int32)
    acc = np.zeros(nT, np.int32)
    for s in range(nsw):
        for k in range(nT):
            one = np.ascontiguousarray(ref[k:k + 1])
            mcref.cb_sweeps(1, 2, nbr, J, S, D, g.colour, [betas[rung[k]]], [0.0], np.array([k], np.uint32), SEED, s,
                            1, one, ising_table=b.ising_table([rung[k]]))
            ref[k] = one[0]
        E = np.array([mcref.cb_energy(1, nbr, J, S, D, 0.0, ref[k])[0] for k in range(nT)])
        mcref.cb_pt(betas, E, who, rung, SEED, 0, s, acc)
    assert np.array_equal(got, ref)
    assert np.array_equal(b.rung.cpu().numpy(), rung)
    assert np.array_equal(b.who.cpu().numpy(), who)
    assert np.array_equal(b.pt_acc.cpu().numpy(), acc)


def test_lds_without_link_patterns(gpu):
    """Irregular graph (no translation patterns): LDS kernel on the global
    link table, still bit-exact."""
    rng = np.random.default_rng(2)
    n, z = 300, 5
    nbr = rng.integers(0, n, (n, z))
    nbr[rng.random((n, z)) < 0.2] = -1
    J = rng.normal(size=(n, z, 9)) * 0.5
    S = np.where(rng.random(n) < 0.5, 1.0, 1.5)
    D = rng.normal(size=(n, 3)) * 0.1
    g = compile_graph(3, nbr, J, S, D)
    assert g.npat == 0
    betas = [0.7, 1.3]
    init = []
    for _ in betas:
        v = rng.normal(size=(n, 3))
        init.append(v / np.linalg.norm(v, axis=1, keepdims=True) * S[:, None])
    b = Batch(3, [g], [ReplicaSpec(0, k) for k in range(2)], betas, [0.2, 0.2], init, seed=SEED)
    b.init()
    b.sweeps(0, 6, kernel=L.KERNEL_LDS)
    got = _gpu_spins(b, 2)
    ref = _cpu_spins(3, init, "f32")
    mcref.cb_sweeps(3, 0, nbr, J, S, D, g.colour, betas, [0.2, 0.2], np.arange(2, dtype=np.uint32), SEED, 0, 6, ref)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


# ---------------------------------------------------------------------------
# replicas split over several workgroups (parts exchange each colour class)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("parts", [2, 4])
@pytest.mark.parametrize("case", sorted(CASES))
def test_split_bitexact(gpu, case, parts):
    """A replica split over 2 or 4 workgroups follows the checker's
    trajectory bit for bit (the Philox groups are dealt out, not changed)."""
    nc, mk, betas, h = CASES[case]
    Lt = mk()
    print("case", case, "parts", parts, "n", Lt.n, flush=True)
    b, g, init = _mk_batch(nc, Lt, betas, h)
    nsw = 12
    b.sweeps(0, nsw, kernel=L.KERNEL_LDS, parts=parts)
    got = _gpu_spins(b, len(betas))
    ref = _cpu_spins(nc, init, "f32")
    nbr, J, S, D = _lattice_tables(Lt, nc)
    mcref.cb_sweeps(nc, 0 if nc > 1 else 2, nbr, J, S, D, g.colour, betas, [h] * len(betas),
                    np.arange(len(betas), dtype=np.uint32), SEED, 0, nsw, ref)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), "spins differ"
    st = b.state_data()
    for k in range(len(betas)):
        e_cpu = mcref.cb_energy(nc, nbr, J, S, D, h, ref[k])
        assert abs(st[k, L.S_E] - e_cpu[0]) <= 1e-3 * max(1.0, abs(e_cpu[0])), (k, st[k, L.S_E], e_cpu[0])
        np.testing.assert_allclose(st[k, L.S_M:L.S_M + 3], e_cpu[1:], rtol=1e-4, atol=1e-3)
    assert b.error_flag() == 0


def _ladder_batch(parts_seed=0):
    """Two tempering ladders (XY L=16 and L=24, 8 rungs each) with every
    measurement on."""
    graphs, reps, betas, spins = [], [], [], []
    rng = np.random.default_rng(11)
    for gi, L_ in enumerate((16, 24)):
        Lt = lat.square(L_, J=-1.0, D=(0.05, 0.0, 0.0))
        graphs.append(compile_graph(2, Lt.nbr, Lt.J, Lt.S, Lt.D, pairs=Lt.corr_pairs, rOrb=Lt.rOrb,
                                    rOrbCluster=Lt.rOrbCluster, rlink=Lt.rlink))
        for r in range(8):
            reps.append(ReplicaSpec(gi, gi * 8, nrung=8, ladder=gi, rung=r))
            betas.append(1 / (0.7 + 0.07 * r))
            v = rng.normal(size=(Lt.n, 2))
            spins.append(v / np.linalg.norm(v, axis=1, keepdims=True))
    return graphs, reps, betas, spins


@pytest.mark.parametrize("parts", [2, 3])
def test_split_run_matches_unsplit(gpu, parts):
    """Measured run with tempering: split and unsplit replicas give the same
    spins, swaps and accumulators (up to FP64 summation order)."""
    out = []
    for pp in (1, parts):
        graphs, reps, betas, spins = _ladder_batch()
        b = Batch(2, graphs, reps, betas, [0.02] * len(betas), spins, meas=L.MEAS_PAIRS | L.MEAS_RNORM, nbins=4,
                  nframes=40, recompute_every=7)
        b.init()
        b.run(10, 40, 1, pt_every=1, kernel=L.KERNEL_LDS, parts=pp)
        out.append((np.concatenate([b.replica_spins(k).ravel() for k in range(len(reps))]), b.accumulators(),
                    b.bin_data(),
                    b.pt_acc.cpu().numpy(), b.rung.cpu().numpy(), b.state_data(), b.error_flag()))
    (s1, a1, bn1, p1, r1, st1, e1), (s2, a2, bn2, p2, r2, st2, e2) = out
    assert e1 == 0 and e2 == 0
    assert np.array_equal(s1.view(np.uint32), s2.view(np.uint32))
    assert np.array_equal(p1, p2) and np.array_equal(r1, r2)
    # A persistent launch with split replicas ends at each FP64 recompute,
    # which the host queues after it, so the frame of a recompute sweep sees
    # the tracked energy there and the recomputed one on unsplit replicas:
    # they differ by the tracking's FP32 rounding (~1e-9 of E here)
    np.testing.assert_allclose(a2, a1, rtol=1e-7, atol=1e-9)
    np.testing.assert_allclose(bn2, bn1, rtol=1e-7, atol=1e-9)
    np.testing.assert_allclose(st2[:, :4], st1[:, :4], rtol=1e-9, atol=1e-9)


def test_auto_split_large_replicas(gpu):
    """Auto planning splits replicas larger than one CU's share (here two
    L=128 XY replicas, 4 parts each) and keeps the trajectory."""
    nc, Lt = 2, lat.square(128, J=-1.0)
    betas = [0.9, 1.4]
    res = []
    for parts in (1, 0):
        b, g, init = _mk_batch(nc, Lt, betas, 0.0)
        b.sweeps(0, 5, kernel=L.KERNEL_LDS, parts=parts)
        res.append(_gpu_spins(b, 2))
        assert b.error_flag() == 0
    ref = _cpu_spins(nc, init, "f32")
    nbr, J, S, D = _lattice_tables(Lt, nc)
    mcref.cb_sweeps(nc, 0, nbr, J, S, D, g.colour, betas, [0.0, 0.0], np.arange(2, dtype=np.uint32), SEED, 0, 5, ref)
    assert np.array_equal(res[0].view(np.uint32), ref.view(np.uint32))
    assert np.array_equal(res[1].view(np.uint32), ref.view(np.uint32))


def _ladder_batch_shared():
    """Three ladders, two of them on one graph (XY L=16: rungs 0-3 and 4-7,
    so that one workgroup holds replicas of both), one on L=24."""
    graphs, reps, betas, spins = [], [], [], []
    rng = np.random.default_rng(12)
    for gi, L_ in enumerate((16, 24)):
        Lt = lat.square(L_, J=-1.0, D=(0.05, 0.0, 0.0))
        graphs.append(compile_graph(2, Lt.nbr, Lt.J, Lt.S, Lt.D, pairs=Lt.corr_pairs))
        for r in range(8):
            if gi == 0:
                reps.append(ReplicaSpec(0, 4 * (r // 4), nrung=4, ladder=r // 4, rung=r % 4))
            else:
                reps.append(ReplicaSpec(1, 8, nrung=8, ladder=2, rung=r))
            betas.append(1 / (0.7 + 0.07 * r))
            v = rng.normal(size=(Lt.n, 2))
            spins.append(v / np.linalg.norm(v, axis=1, keepdims=True))
    return graphs, reps, betas, spins


def _run_ladders(parts, no_persist, pt_every, nth=10, nsw=40, meas=0, build=None):
    graphs, reps, betas, spins = (build or _ladder_batch)()
    b = Batch(2, graphs, reps, betas, [0.02] * len(betas), spins, meas=meas, nbins=4, nframes=nsw)
    b.init()
    b.run(nth, nsw, 1, pt_every=pt_every, kernel=L.KERNEL_LDS, parts=parts, no_persist=no_persist)
    return (np.concatenate([b.replica_spins(k).ravel() for k in range(len(reps))]), b.accumulators(), b.bin_data(),
            b.pt_acc.cpu().numpy(), b.rung.cpu().numpy(), b.who.cpu().numpy(), b.state_data(), b.error_flag())


@pytest.mark.parametrize("parts,pt_every", [(1, 1), (1, 3), (2, 1), (2, 0), (3, 2)])
def test_persistent_matches_per_launch(gpu, parts, pt_every):
    """A persistent launch (grid barrier, tempering and part exchange inside
    the kernel) runs the same trajectory as one launch per sweep."""
    a = _run_ladders(parts, True, pt_every)
    b = _run_ladders(parts, False, pt_every)
    assert a[-1] == 0 and b[-1] == 0
    assert np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32))
    for x, y in zip(a[3:6], b[3:6]):
        assert np.array_equal(x, y)
    np.testing.assert_allclose(b[1], a[1], rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(b[2], a[2], rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(b[6][:, :4], a[6][:, :4], rtol=1e-11, atol=1e-9)


@pytest.mark.parametrize("parts", [1, 2])
def test_persistent_shared_graph_ladders(gpu, parts):
    """Ladders that share a graph (and workgroups): the persistent launch
    takes each workgroup's ladders from its replicas and writes every
    ladder back once."""
    a = _run_ladders(parts, True, 1, build=_ladder_batch_shared)
    b = _run_ladders(parts, False, 1, build=_ladder_batch_shared)
    assert a[-1] == 0 and b[-1] == 0
    assert np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32))
    for x, y in zip(a[3:6], b[3:6]):
        assert np.array_equal(x, y)
    assert a[3].sum() > 0
    np.testing.assert_allclose(b[1], a[1], rtol=1e-11, atol=1e-11)


def _honeycomb_ladders():
    """Heisenberg honeycomb (full tensors, Dz, one triangle per cell, two
    orbital groups), two 6-rung ladders of L = 16 and 24 cells: every
    spin-derived observable (pairs, Q, block spins, groups) is measured."""
    from cases import _honeycomb
    graphs, reps, betas, spins = [], [], [], []
    rng = np.random.default_rng(21)
    for gi, L_ in enumerate((16, 24)):
        Lt = _honeycomb(L_, 5 + gi)
        graphs.append(compile_graph(3, Lt.nbr, Lt.J, Lt.S, Lt.D, pairs=Lt.corr_pairs, circuits=Lt.circuits,
                                    rOrb=Lt.rOrb, rOrbCluster=Lt.rOrbCluster, rlink=Lt.rlink, groups=Lt.groups))
        for r in range(6):
            reps.append(ReplicaSpec(gi, gi * 6, nrung=6, ladder=gi, rung=r))
            betas.append(1 / (0.6 + 0.1 * r))
            v = rng.normal(size=(Lt.n, 3))
            spins.append(1.5 * v / np.linalg.norm(v, axis=1, keepdims=True))
    return graphs, reps, betas, spins


def _honeycomb_ladders_local():
    """Two 6-rung ladders (L = 16 and 24 cells) of a Heisenberg honeycomb
    with the default correlation pairs (orbital 0 of every cell with
    itself: half the sites, so the sums are taken from the spins) and no
    circuits: no pair has sites of two parts."""
    from cases import _honeycomb
    graphs, reps, betas, spins = [], [], [], []
    rng = np.random.default_rng(23)
    for gi, L_ in enumerate((16, 24)):
        Lt = _honeycomb(L_, 5 + gi)
        p = Lt.corr_pairs
        assert (p[:, 0] == p[:, 1]).all() and len(p) < Lt.n  # summed from the spins
        graphs.append(compile_graph(3, Lt.nbr, Lt.J, Lt.S, Lt.D, pairs=p))
        for r in range(6):
            reps.append(ReplicaSpec(gi, gi * 6, nrung=6, ladder=gi, rung=r))
            betas.append(1 / (0.6 + 0.1 * r))
            v = rng.normal(size=(Lt.n, 3))
            spins.append(1.5 * v / np.linalg.norm(v, axis=1, keepdims=True))
    return graphs, reps, betas, spins


def _honeycomb_ladders_cross_pairs():
    """Two 6-rung ladders (L = 16 and 24 cells) of a Heisenberg honeycomb
    whose correlation pairs join orbital 0 of a cell to orbital 1 of the cell
    three rows on (ki_s = 0, ki_t = 1, overLat (0, 3, 0)), and one triangle
    per cell: pairs and triangles with sites of two parts."""
    graphs, reps, betas, spins = [], [], [], []
    rng = np.random.default_rng(22)
    for gi, L_ in enumerate((16, 24)):
        r = np.random.default_rng(5 + gi)
        bonds = []
        for (s_, t, o) in [(0, 1, (0, 0, 0)), (0, 1, (-1, 0, 0)), (0, 1, (0, 1, 0)), (0, 0, (1, 0, 0)),
                           (1, 1, (1, 0, 0))]:
            J = np.zeros(9)
            J[:3] = -1 + r.uniform(-0.1, 0.1, 3)
            J[3:] = r.uniform(-0.05, 0.05, 6)
            bonds.append(lat.Bond(s_, t, o, J))
        Lt = lat.build_lattice(L_, L_, 1, [[1, 0, 0], [-0.5, np.sqrt(3) / 2, 0], [0, 0, 1]],
                               [[1 / 3, 2 / 3, 0], [2 / 3, 1 / 3, 0]], bonds, [1.5, 1.5], [[0, 0, -0.1]] * 2,
                               ki_s=0, ki_t=1, ki_overLat=(0, 3, 0),
                               localCircuitList=[((0, (0, 0, 0)), (1, (0, 0, 0)), (0, (1, 0, 0)))])
        assert (Lt.corr_pairs[:, 0] != Lt.corr_pairs[:, 1]).all()
        graphs.append(compile_graph(3, Lt.nbr, Lt.J, Lt.S, Lt.D, pairs=Lt.corr_pairs, circuits=Lt.circuits))
        for k in range(6):
            reps.append(ReplicaSpec(gi, gi * 6, nrung=6, ladder=gi, rung=k))
            betas.append(1 / (0.6 + 0.1 * k))
            v = rng.normal(size=(Lt.n, 3))
            spins.append(1.5 * v / np.linalg.norm(v, axis=1, keepdims=True))
    return graphs, reps, betas, spins


def _run_full(build, nc, parts, no_persist, pt_every, meas, nth=10, nsw=30):
    graphs, reps, betas, spins = build()
    # no FP64 recompute: it resets the tracked energy at different points of
    # a split and an unsplit sweep (see test_split_run_matches_unsplit)
    b = Batch(nc, graphs, reps, betas, [0.02] * len(betas), spins, meas=meas, nbins=4, nframes=nsw,
              recompute_every=0)
    b.init()
    b.run(nth, nsw, 1, pt_every=pt_every, kernel=L.KERNEL_LDS, parts=parts, no_persist=no_persist)
    return (np.concatenate([b.replica_spins(k).ravel() for k in range(len(reps))]).astype(np.float32),
            b.accumulators(), b.bin_data(), b.pt_acc.cpu().numpy(), b.rung.cpu().numpy(), b.who.cpu().numpy(),
            b.state_data(), b.gacc.cpu().numpy() if b.gacc is not None else None, b.error_flag())


@pytest.mark.parametrize("model,parts,pt_every", [("xy", 1, 1), ("xy", 2, 1), ("xy", 3, 0), ("xy", 2, 3),
                                                  ("heis", 1, 1), ("heis", 2, 1), ("heis", 4, 0),
                                                  ("xy_pairs", 2, 1), ("heis_pq", 2, 1), ("heis_pq", 3, 0),
                                                  ("heis_pq", 4, 1), ("heis_local", 2, 1), ("heis_local", 3, 0),
                                                  ("heis_local", 4, 1)])
def test_persistent_spin_observables_match_per_launch(gpu, model, parts, pt_every):
    """Spin-derived observables (pair correlations, block spins; for the
    Heisenberg honeycomb also Q and orbital groups) measured inside the
    persistent launch: unsplit replicas from their LDS copies; split ones,
    when no pair or circuit has sites of two parts and no block spins or
    groups are measured (xy_pairs, heis_local), each part over the pairs
    and circuits whose first site it owns, from its own copy before the grid
    barrier, otherwise (heis, heis_pq with pairs and triangles across
    parts) from every part's published sites, each part summing its share;
    part 0 adds the parts' sums in part order.  Same spins, swaps and rungs
    bit for bit as one launch per sweep, accumulators, bins and group sums
    to FP64 summation order."""
    if model == "xy":
        build, nc, meas = _ladder_batch, 2, L.MEAS_PAIRS | L.MEAS_RNORM
    elif model == "xy_pairs":
        build, nc, meas = _ladder_batch, 2, L.MEAS_PAIRS
    elif model == "heis_pq":
        build, nc, meas = _honeycomb_ladders_cross_pairs, 3, L.MEAS_PAIRS | L.MEAS_Q
    elif model == "heis_local":
        build, nc, meas = _honeycomb_ladders_local, 3, L.MEAS_PAIRS
    else:
        build, nc, meas = _honeycomb_ladders, 3, L.MEAS_PAIRS | L.MEAS_RNORM | L.MEAS_Q | L.MEAS_GROUPS
    a = _run_full(build, nc, parts, True, pt_every, meas)
    b = _run_full(build, nc, parts, False, pt_every, meas)
    assert a[-1] == 0 and b[-1] == 0
    assert np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32))
    for x, y in zip(a[3:6], b[3:6]):
        assert np.array_equal(x, y)
    np.testing.assert_allclose(b[1], a[1], rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(b[2], a[2], rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(b[6][:, :4], a[6][:, :4], rtol=1e-10, atol=1e-8)
    if a[7] is not None:
        np.testing.assert_allclose(b[7], a[7], rtol=1e-10, atol=1e-10)
    # the observables were measured: pair correlations and (Heisenberg) Q
    assert np.abs(a[1][:, L.A_SIJ]).max() > 0
    if model in ("heis", "heis_pq"):
        assert np.abs(a[1][:, L.A_Q]).max() > 0


EXACT_ISING = {
    # J = 1, h = 0 (the Ising 32^2 benchmark's couplings) and a dyadic field
    "sq16_h0": (lambda: lat.square(16, ising=True), [0.2, 0.44, 0.6, 1.0], 0.0),
    "sq16_h025": (lambda: lat.square(16, J=-0.75, ising=True), [0.3, 0.5], 0.25),
    # two orbitals with spin lengths 1 and 1.5, equal couplings
    "2orb_S15": (lambda: __import__("test_gpu_dropin")._ising_2orb_uniform(6), [0.3, 0.7], 0.5),
}


@pytest.mark.parametrize("kernel", [L.KERNEL_LDS, L.KERNEL_STREAM, L.KERNEL_PLAIN])
@pytest.mark.parametrize("case", sorted(EXACT_ISING))
def test_exact_ising_bitexact(gpu, case, kernel):
    """Dyadic couplings, fields and spin lengths: the sweep accepts through
    the uint32 threshold table (exact_ising_table), and so does the checker;
    spins bit for bit, and the tracked energy equals the recomputed one
    exactly (every dE is an exact multiple of the unit)."""
    mk, betas, h = EXACT_ISING[case]
    Lt = mk()
    b, g, init = _mk_batch(1, Lt, betas, h)
    tab = b.ising_table()
    assert tab is not None, "expected the exact acceptance"
    nsw = 12
    b.sweeps(0, nsw, kernel=kernel)
    got = _gpu_spins(b, len(betas))
    ref = _cpu_spins(1, init, "f32")
    nbr, J, S, D = _lattice_tables(Lt, 1)
    mcref.cb_sweeps(1, 2, nbr, J, S, D, g.colour, betas, [h] * len(betas), np.arange(len(betas), dtype=np.uint32),
                    SEED, 0, nsw, ref, ising_table=tab)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), "spins differ"
    st = b.state_data()
    for k in range(len(betas)):
        e_cpu = mcref.cb_energy(1, nbr, J, S, D, h, ref[k])
        assert st[k, L.S_E] == e_cpu[0], (k, st[k, L.S_E], e_cpu[0])
    # the sweep reads the table: with every threshold 0 (the T = 0 limit:
    # only flips with dE <= 0) it follows the checker given the same table
    b.ithr.zero_()
    b.sweeps(nsw, nsw, kernel=kernel)
    got = _gpu_spins(b, len(betas))
    zero = (np.zeros_like(tab[0]), tab[1])
    mcref.cb_sweeps(1, 2, nbr, J, S, D, g.colour, betas, [h] * len(betas), np.arange(len(betas), dtype=np.uint32),
                    SEED, nsw, nsw, ref, ising_table=zero)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), "spins differ with the zero table"
    for k in range(len(betas)):
        assert mcref.cb_energy(1, nbr, J, S, D, h, ref[k])[0] <= b.state_data()[k, L.S_E] + 0.0


@pytest.mark.parametrize("block", [256, 512])
@pytest.mark.parametrize("case", ["ising_sq16", "xy_sq16", "xy_sq24_pool", "heis_honeycomb"])
def test_block_sizes_bitexact(gpu, case, block):
    """The LDS kernel with 256- and 512-thread workgroups (an autotuner
    dimension) follows the checker bit for bit."""
    nc, mk, betas, h = CASES[case]
    Lt = mk()
    print("case", case, "block", block, flush=True)
    b, g, init = _mk_batch(nc, Lt, betas, h)
    nsw = 10
    b.sweeps(0, nsw, kernel=L.KERNEL_LDS, block=block)
    got = _gpu_spins(b, len(betas))
    ref = _cpu_spins(nc, init, "f32")
    nbr, J, S, D = _lattice_tables(Lt, nc)
    mcref.cb_sweeps(nc, 0 if nc > 1 else 2, nbr, J, S, D, g.colour, betas, [h] * len(betas),
                    np.arange(len(betas), dtype=np.uint32), SEED, 0, nsw, ref, ising_table=b.ising_table())
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), "spins differ"
    assert b.error_flag() == 0


@pytest.mark.parametrize("block,parts", [(256, 1), (512, 1), (256, 2), (512, 3)])
def test_block_sizes_persistent_ladders(gpu, block, parts):
    """Persistent launches (tempering every sweep) with 256- and 512-thread
    workgroups, split or not: same spins, rungs and swaps as the planner's
    1024-thread plan."""
    def run(blk):
        graphs, reps, betas, spins = _ladder_batch()
        b = Batch(2, graphs, reps, betas, [0.02] * len(betas), spins, nbins=4, nframes=30)
        b.init()
        b.run(5, 30, 1, pt_every=1, kernel=L.KERNEL_LDS, block=blk, parts=parts)
        return (np.concatenate([b.replica_spins(k).ravel() for k in range(len(reps))]).astype(np.float32),
                b.rung.cpu().numpy(), b.who.cpu().numpy(), b.pt_acc.cpu().numpy(), b.error_flag())
    a, c = run(block), run(0)
    assert a[-1] == 0 and c[-1] == 0
    assert np.array_equal(a[0].view(np.uint32), c[0].view(np.uint32))
    for x, y in zip(a[1:4], c[1:4]):
        assert np.array_equal(x, y)
