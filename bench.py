"""Benchmark: spin updates/s of the batched MI355X engine on the XY
Binder-scan workload (7 square lattices L = 16..128 mixed in one launch, each
a 64-rung parallel-tempering ladder over T in [0.7, 1.2], swap and measure
every sweep, FP32 spins, FP64 accumulators).

One step = one sweep of every replica (2,166,784 spin updates) + one
tempering swap pass + one measured frame.  Printed as one JSON line with:
  value        whole-job spin updates/s (inputs resident in HBM), timed with
               two events around the K steps
  roofline     dominant kernel (k_lds) vs the 8 TB/s HBM peak: a second pass
               of K/4 steps stamps every launch of it with HIP events (from
               its own dispatch), and 2 x (state bytes) per site update is
               the algorithmic traffic; `traffic` is the HBM bytes per
               launch that rocprofv3 PMC counters measured for this kernel
               and workload (committed summary under profiles/)
  cpu_baseline the FP64 reference-algorithm restatement (oracle), random
               sequential Metropolis with glibc rand(), on the host cores
  plain        the same trajectory with one launch per replica per colour per
               sweep (global-memory stencil) and the resulting speedup

python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--no-plain]
               [--workload NAME] [--store f32|f16] [--parts P]

--workload selects one of the other workloads (ising_square32: Ising 32^2 x
16 T; heis_honeycomb64: Heisenberg honeycomb 64x64x2 x 64 T;
heis_triangular_dmi128: triangular DMI 128^2 x 1024 (T, B); heis_cubic64:
cubic 64^3 x 128 T) and prints the same fields for it (without the plain and
CPU legs); the default is the XY ladder scan above.

--workload cubic_wolff_vs_metropolis compares Wolff cluster updates with
batched Metropolis on the cubic 64^3 Heisenberg ferromagnet (16 independent
replicas at --tau-T each): frames/s, tau_int of |M| and E from per-frame
series, and independent samples/s = replicas x frames/s / (2 tau_int).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "spin updates/s (replica·sites/s) + % HBM roofline; speedup vs plain at 2σ U4"
HBM_PEAK = 8.0e12
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_lds_xy_scan_hbm_traffic.md")


def profiled_traffic(sweeps_per_launch: float):
    """HBM bytes of one launch of the XY-scan sweep kernel running the given
    number of sweeps, from rocprofv3 PMC counters (FETCH_SIZE, WRITE_SIZE,
    each scaled by a copy-kernel calibration) in the committed profile
    summary: a fixed part per launch plus a part per sweep."""
    vals = {}
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                for key in ("traffic_fixed_bytes_per_launch:", "traffic_bytes_per_sweep:"):
                    if line.startswith(key):
                        vals[key] = float(line.split(":")[1])
    except OSError:
        return None
    if len(vals) != 2:
        return None
    return vals["traffic_fixed_bytes_per_launch:"] + sweeps_per_launch * vals["traffic_bytes_per_sweep:"]


def _cpu_worker(args):
    L, T, nsweep, seed = args
    from mcsolver_amd import lattice as lat
    from oracle import mcref
    Lt = lat.square(L, J=-1.0)
    a = lat.on_args(Lt, T, 0.0, 0, nsweep)
    t0 = time.perf_counter()
    mcref.on_run(2, a, seed=seed)
    return Lt.n * nsweep, time.perf_counter() - t0


def cpu_baseline(target_s: float = 10.0):
    """Reference algorithm (oracle restatement of xyLib.c MCMainFunction,
    random-sequential Metropolis + full observable set, FP64, glibc rand())
    on L=32 points of the same T ladder, one process per point as
    mcsolver/win.py:81-103 pumps them."""
    P = min(16, len(os.sched_getaffinity(0)))
    # calibrate one point on this core
    upd, dt = _cpu_worker((32, 0.9, 200, 1))
    rate1 = upd / dt
    nsweep = max(200, int(target_s * rate1 / 1024))
    temps = [0.7 + 0.5 * k / (P - 1 if P > 1 else 1) for k in range(P)]
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(P) as pool:
        res = pool.map(_cpu_worker, [(32, T, nsweep, 100 + k) for k, T in enumerate(temps)])
    wall = time.perf_counter() - t0
    tot = sum(r[0] for r in res)
    return {"value": tot / wall, "unit": "spin updates/s", "cores": P, "kind": "port",
            "sample": "XY L=32, %d temperatures of the [0.7,1.2] ladder, %d sweeps each, FP64 oracle restatement of "
                      "xyLib.c MCMainFunction (random-sequential Metropolis, glibc rand, full observables), one "
                      "process per point; %.1f s wall" % (P, nsweep, wall),
            "per_core": tot / sum(r[1] for r in res)}


def copy_calibration(mib: int) -> None:
    """One device-to-device copy of `mib` MiB: a kernel of known traffic, for
    scaling rocprofv3's FETCH_SIZE / WRITE_SIZE counters."""
    if mib > 0:
        import torch
        src = torch.empty(mib << 20, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst


def _energy_check(b):
    """The run's full FP64 recomputes: their period, and the largest
    |E_tracked - E_full| any replica showed (absolute, and per |E|)."""
    from mcsolver_amd import _lib as L
    st = b.state_data()
    d = st[:, L.S_DRIFT]
    rel = d / np.maximum(np.abs(st[:, L.S_E]), 1e-300)
    return {"recompute_every": int(b.b.recompute_every), "max_abs_drift": float(d.max()),
            "max_rel_drift": float(rel.max()), "error_flag": b.error_flag()}


def tuned_options(a, b, pt_every):
    """Plan options for Batch.run: the command line's, or with --autotune the
    fastest of the candidate plans timed on this batch (the batch is left as
    it was)."""
    if not a.autotune:
        return dict(parts=a.parts, no_persist=False, reps_per_block=0)
    from mcsolver_amd.autotune import autotune
    r = autotune(b, 1, pt_every, trial_sweeps=64)
    for t in r.trials:
        print("autotune %s: %s" % (t.options, "%.4f ms/sweep" % t.ms_per_sweep if t.ms_per_sweep is not None
                                    else "rejected (%s)" % t.error), file=sys.stderr, flush=True)
    return r.best


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--plain-steps", type=int, default=20)
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (HBM counter calibration)")
    ap.add_argument("--workload", default="xy_ladder_scan",
                    choices=["xy_ladder_scan", "ising_square32", "heis_honeycomb64", "heis_triangular_dmi128",
                             "heis_cubic64", "cubic_wolff_vs_metropolis"])
    ap.add_argument("--tau-T", type=float, default=1.44,
                    help="cubic_wolff_vs_metropolis: temperature (T_c of the 3D Heisenberg ferromagnet is ~1.443)")
    ap.add_argument("--tau-seconds", type=float, default=40.0,
                    help="cubic_wolff_vs_metropolis: GPU time per method for the measured series")
    ap.add_argument("--wolff-clusters", type=int, default=32,
                    help="cubic_wolff_vs_metropolis: Wolff clusters per replica between frames")
    ap.add_argument("--store", default="f32", choices=["f32", "f16"])
    ap.add_argument("--parts", type=int, default=0,
                    help="workgroups per LDS replica: 0 = balance automatically, 1 = never split")
    ap.add_argument("--autotune", action="store_true",
                    help="choose the launch plan by timing candidate plans on the batch first (outside the timed "
                         "region; results do not depend on the plan)")
    a = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print("bench.py runs on one GPU only: replicas shard across GPUs with no communication during sweeps, "
              "so N>1 is the same kernel on 1/N of the batch; multi-GPU runs are not part of this build.",
              file=sys.stderr)
        sys.exit(2)
    if a.workload == "cubic_wolff_vs_metropolis":
        wolff_vs_metropolis(a)
        return
    if a.workload != "xy_ladder_scan":
        other_workload(a)
        return
    # CPU baseline first (forks before the GPU is initialised)
    cpu = None if a.no_cpu else cpu_baseline()

    import torch
    from mcsolver_amd import _lib as L
    from mcsolver_amd import workloads as W

    torch.cuda.set_device(0)
    copy_calibration(a.copy_calib_mib)
    K, Wm = a.steps, a.warmup
    KT = max(1, K // 4)  # sweeps of the separate per-launch timing pass
    wl = W.build("xy_ladder_scan", nframes=K + Wm + KT, nbins=16)
    b = wl.batch
    b.init()
    opts = tuned_options(a, b, wl.pt_every)
    b.run(0, Wm, 1, pt_every=wl.pt_every, block=a.block, **opts)
    torch.cuda.synchronize()

    def plan(nsw, sweep0, timing=None):
        return b.plan(0, nsw, 1, L.METROPOLIS, wl.pt_every, L.KERNEL_AUTO, a.block, opts["reps_per_block"], 0,
                      sweep0, sweep0, opts["parts"], opts["no_persist"], timing)

    # timed region: K steps, whole job, no per-launch events.  The warmup ran
    # the same plan shape, so its prepared run (launch plan, tables, device
    # scratch) is reused: the timed call only queues kernel launches.
    st, en = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    p = plan(K, Wm)
    b.runner(p)
    torch.cuda.synchronize()
    st.record()
    b.run_plan(p)
    en.record()
    torch.cuda.synchronize()
    ms = st.elapsed_time(en)
    ups = wl.updates_per_sweep
    value = ups * K / (ms / 1e3)
    # roofline pass: the same steps with HIP events around every launch of the
    # sweep kernel on the run's stream (the events cost a little wall time,
    # so this pass is kept out of `value`)
    timing = (C.c_double * 4)()
    b.run_plan(plan(KT, Wm + K, timing))
    torch.cuda.synchronize()
    bpu = W.bytes_per_update(2, "f32")
    k_ms, k_launch, k_sweeps, k_upd = timing[0], timing[1], timing[2], timing[3]
    achieved = (k_upd * bpu) / (k_ms / 1e3) if k_ms > 0 else 0.0
    roof = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": achieved / HBM_PEAK, "traffic": profiled_traffic(k_sweeps / max(k_launch, 1)),
            "kernel": "k_lds<2,float>",
            "avg_launch_us": 1e3 * k_ms / max(k_launch, 1), "launches": int(k_launch),
            "algorithmic_bytes_per_launch": bpu * k_upd / max(k_launch, 1)}
    acc = b.accumulators()
    assert b.error_flag() == 0
    assert (acc[:, L.A_COUNT] == K + Wm + KT).all(), "every temperature row must have one frame per sweep"
    energy_check = _energy_check(b)
    out = {"metric": METRIC, "value": value, "unit": "spin updates/s", "n_gpus": 1, "steps": K, "warmup": Wm,
           "ms_per_step": ms / K, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic",
           "config": {"workload": "XY Binder scan, 7 square lattices x 64-rung PT ladders", "lattices": list(W.XY_SCAN_SIZES),
                      "rungs_per_ladder": 64, "replicas": len(b.reps), "spins_per_sweep": ups, "pt_every": 1,
                      "measure_every": 1, "spin_storage": "fp32", "accumulators": "fp64"},
           "roofline": roof, "cpu_baseline": cpu, "energy_check": energy_check}
    if a.autotune:
        out["plan"] = opts
    if not a.no_plain:
        wp = W.build("xy_ladder_scan", nframes=a.plain_steps, nbins=1)
        wp.batch.init()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        wp.batch.run(0, a.plain_steps, 1, pt_every=wp.pt_every, kernel=L.KERNEL_PLAIN)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        pv = wp.updates_per_sweep * a.plain_steps / dt
        out["plain"] = {"value": pv, "unit": "spin updates/s", "steps": a.plain_steps,
                        "kernel": "k_pass, one launch per replica per colour per sweep + per-replica measure"}
        out["speedup_vs_plain"] = value / pv
    print(json.dumps(out), flush=True)


def other_workload(a):
    """One of the other workloads: K measured sweeps of every replica (no
    tempering, so one launch runs all sweeps on the LDS path), then a K/4
    sweep pass with per-launch HIP events for the roofline."""
    import torch
    from mcsolver_amd import _lib as L
    from mcsolver_amd import workloads as W

    torch.cuda.set_device(0)
    copy_calibration(a.copy_calib_mib)
    K, Wm = a.steps, a.warmup
    KT = max(1, K // 4)
    meas = L.MEAS_PAIRS | (L.MEAS_Q if a.workload == "heis_triangular_dmi128" else 0)
    wl = W.build(a.workload, store=a.store, nframes=K + Wm + KT, nbins=16, meas=meas)
    b = wl.batch
    b.init()
    opts = tuned_options(a, b, 0)
    b.run(0, Wm, 1, block=a.block, **opts)
    torch.cuda.synchronize()

    def plan(nsw, sweep0, timing=None):
        return b.plan(0, nsw, 1, L.METROPOLIS, 0, L.KERNEL_AUTO, a.block, opts["reps_per_block"], 0, sweep0,
                      sweep0, opts["parts"], opts["no_persist"], timing)

    st, en = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    p = plan(K, Wm)
    b.runner(p)
    torch.cuda.synchronize()
    st.record()
    b.run_plan(p)
    en.record()
    torch.cuda.synchronize()
    ms = st.elapsed_time(en)
    ups = wl.updates_per_sweep
    value = ups * K / (ms / 1e3)
    timing = (C.c_double * 4)()
    b.run_plan(plan(KT, Wm + K, timing))
    torch.cuda.synchronize()
    bpu = W.bytes_per_update(wl.nc, wl.store)
    k_ms, k_launch, k_sweeps, k_upd = timing[0], timing[1], timing[2], timing[3]
    stream = k_launch > k_sweeps  # one launch per colour: the streaming pass
    ty = {"f32": "float", "f16": "__half", "i8": "signed char"}[wl.store]
    kern = ("k_pass<%d,%s>" if stream else "k_lds<%d,%s>") % (wl.nc, ty)
    achieved = (k_upd * bpu) / (k_ms / 1e3) if k_ms > 0 else 0.0
    acc = b.accumulators()
    assert b.error_flag() == 0
    assert (acc[:, L.A_COUNT] == K + Wm + KT).all()
    energy_check = _energy_check(b)
    out = {"metric": METRIC, "value": value, "unit": "spin updates/s", "n_gpus": 1, "steps": K, "warmup": Wm,
           "ms_per_step": ms / K, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": {"f32": "f32", "f16": "f32 (f16 storage)", "i8": "f32 (int8 spins)"}[wl.store],
           "data": "synthetic",
           "config": {"workload": wl.name + ": " + wl.description, "replicas": len(b.reps), "spins_per_sweep": ups,
                      "measure_every": 1, "spin_storage": wl.store, "accumulators": "fp64"},
           "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK, "traffic": None, "kernel": kern,
                        "avg_launch_us": 1e3 * k_ms / max(k_launch, 1), "launches": int(k_launch),
                        "algorithmic_bytes_per_launch": bpu * k_upd / max(k_launch, 1)},
           "cpu_baseline": None, "energy_check": energy_check}
    if a.autotune:
        out["plan"] = opts
    print(json.dumps(out), flush=True)


def _series(wl, algorithm, ncl, pre_metropolis, nthermal, chunk, fmax, budget_s, label):
    """Thermalise (pre_metropolis Metropolis sweeps, then nthermal steps of the
    method), then measure one frame per step in chunks until fmax frames or
    budget_s seconds of GPU time.  Returns (frames, ms, bins[nrow, frames, q],
    spins flipped per replica during the measured steps)."""
    import torch
    from mcsolver_amd import _lib as L
    b = wl.batch
    b.init()
    if pre_metropolis:
        b.run(pre_metropolis, 0, 1, algorithm=L.METROPOLIS)
    b.run(nthermal, 0, 1, algorithm=algorithm, clusters_per_sweep=ncl, sweep0=pre_metropolis)
    torch.cuda.synchronize()
    flips0 = b.state_data()[:, L.S_ACC].copy()
    s0 = pre_metropolis + nthermal
    done, ms = 0, 0.0
    while done < fmax and ms < budget_s * 1e3:
        n = min(chunk, fmax - done)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        b.run(0, n, 1, algorithm=algorithm, clusters_per_sweep=ncl, sweep0=s0 + done, frame0=done)
        e1.record()
        e1.synchronize()
        ms += e0.elapsed_time(e1)
        done += n
        print("%s: %d frames, %.1f s" % (label, done, ms / 1e3), file=sys.stderr, flush=True)
    assert b.error_flag() == 0
    return done, ms, b.bin_data()[:, :done, :], b.state_data()[:, L.S_ACC] - flips0


def wolff_vs_metropolis(a):
    """The cubic 64^3 Heisenberg comparison: Wolff single-cluster updates (one
    workgroup per replica) against batched Metropolis, 16 independent replicas
    each at one temperature, by tau_int-corrected throughput.  For each
    replica tau_int of |M| and of E is estimated from its per-frame series;
    independent samples/s = replicas x frames/s / (2 tau_int)."""
    import numpy as np
    import torch
    from mcsolver_amd import _lib as L
    from mcsolver_amd import workloads as W
    from mcsolver_amd.analysis import tau_int

    torch.cuda.set_device(0)
    nrep, T = 16, a.tau_T
    out = {"metric": "independent samples/s (tau_int-corrected), Wolff vs batched Metropolis",
           "config": {"workload": "Heisenberg cubic 64^3, J=-1, random start", "T": T, "replicas": nrep,
                      "wolff_clusters_per_frame": a.wolff_clusters, "gpu_seconds_per_method": a.tau_seconds},
           "data": "synthetic"}
    # Wolff starts from a Metropolis-equilibrated state: from a random start
    # its clusters stay small and it orders slowly
    runs = (("metropolis", L.METROPOLIS, 0, 0, 10_000, 5000, 400_000),
            ("wolff", L.WOLFF, a.wolff_clusters, 10_000, 50, 25, 40_000))
    for name, alg, ncl, pre, nth, chunk, fmax in runs:
        wl = W.cubic_replicas(T, nrep, wolff=(alg == L.WOLFF), nframes=fmax)
        n_sites = wl.batch.graphs[0].n
        frames, ms, bins, flips = _series(wl, alg, ncl, pre, nth, chunk, fmax, a.tau_seconds, name)
        fps = frames / (ms / 1e3)
        # sweep-equivalents per frame: one sweep (Metropolis), spins flipped / N (Wolff)
        sweq = 1.0 if alg == L.METROPOLIS else float(flips.mean()) / (n_sites * frames)
        res = {"frames": frames, "seconds": ms / 1e3, "frames_per_s": fps, "sweep_equivalents_per_frame": sweq}
        if alg == L.WOLFF:
            res["mean_flipped_cluster"] = float(flips.mean()) / (frames * ncl)
        for q, obs in ((2, "M"), (0, "E")):
            taus = [tau_int(bins[r, :, q]) for r in range(nrep)]
            tv = np.array([t for t, _ in taus])
            closed = sum(1 for _, w in taus if w < frames - 1)
            res["tau_" + obs] = float(tv.mean())
            res["tau_%s_sweeps" % obs] = float(tv.mean()) * sweq
            res["tau_%s_sem" % obs] = float(tv.std(ddof=1) / np.sqrt(nrep))
            res["windows_closed_" + obs] = closed
            res["frames_over_tau_" + obs] = frames / float(tv.mean())
            res["samples_per_s_" + obs] = nrep * fps / (2.0 * float(tv.mean()))
        res["mean_abs_M"] = float(bins[:, :, 2].mean())
        if alg == L.METROPOLIS:
            res["spin_updates_per_s"] = fps * wl.updates_per_sweep
        out[name] = res
        del wl
        torch.cuda.empty_cache()
    for obs in ("M", "E"):
        out["wolff_over_metropolis_" + obs] = out["wolff"]["samples_per_s_" + obs] / out["metropolis"]["samples_per_s_" + obs]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
