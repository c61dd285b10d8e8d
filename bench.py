"""Benchmark: spin updates/s of the batched MI355X engine on the XY
Binder-scan workload (7 square lattices L = 16..128 mixed in one launch, each
a 64-rung parallel-tempering ladder over T in [0.7, 1.2], swap and measure
every sweep, FP32 spins, FP64 accumulators).

One step = one sweep of every replica (2,166,784 spin updates) + one
tempering swap pass + one measured frame.  Printed as one JSON line with:
  value        whole-job spin updates/s (inputs resident in HBM), timed with
               two events around the K steps
  roofline     the dominant sweep kernel: a separate pass of TIMING_SWEEPS
               sweeps (one launch of the LDS kernel) stamps every launch of it
               with HIP events on the run's stream; `achieved` is its
               algorithmic bytes (2 x spin state per site update) per second
               against the 8 TB/s HBM peak.  From the committed rocprofv3
               summary of the same kernel, workload and launch
               (profiles/<kernel>_<workload>_<storage>.md): `traffic` (HBM
               bytes per launch, FETCH_SIZE + WRITE_SIZE scaled by a copy
               kernel), `traffic_frac` (that traffic per live launch time
               against the peak), `valu_frac` (rocprofv3's VALUBusy) and
               `valu_issue_frac` (VALU instructions per CU-cycle)
  cpu_baseline the same workload's lattice on the host cores, one process per
               temperature point as mcsolver/win.py:81-103 runs them: the
               reference's own C engine compiled from its sources where
               oracle/_ref holds it, else the FP64 restatement (oracle)
  plain        (XY scan) the same trajectory with one launch per replica per
               colour per sweep (global-memory stencil) and the speedup

python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--no-plain]
               [--workload NAME] [--store f32|f16] [--parts P]

--workload selects one of the other workloads (ising_square32: Ising 32^2 x
16 T; heis_honeycomb64: Heisenberg honeycomb 64x64x2 x 64 T;
heis_triangular_dmi128: triangular DMI 128^2 x 1024 (T, B); heis_cubic64:
cubic 64^3 x 128 T) and prints the same fields for it (without the plain
leg); the default is the XY ladder scan above.

--workload cubic_wolff_vs_metropolis compares Wolff cluster updates with
batched Metropolis on the cubic 64^3 Heisenberg ferromagnet (16 independent
replicas at --tau-T each): frames/s, tau_int of |M| and E from per-frame
series, and independent samples/s = replicas x frames/s / (2 tau_int).

python bench.py --profile-summary DIR --workload NAME [--store S] prints the
markdown summary of a profiling directory (DIR/kt, DIR/fetch, DIR/write,
DIR/valu, optionally DIR/mix: rocprofv3 CSV output of the commands listed in
DIR/commands.txt; DIR/bench.json: the profiled run's own JSON line).
"""
from __future__ import annotations

import argparse
import csv
import ctypes as C
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "spin updates/s (replica·sites/s) + % HBM roofline; speedup vs plain at 2σ U4"
HBM_PEAK = 8.0e12
NCU, NXCD = 256, 8
TIMING_SWEEPS = 400   # sweeps of the roofline timing pass
PROFILES = os.path.join(ROOT, "profiles")
FIGURE_KEYS = ("kernel", "sweeps_per_launch", "avg_launch_us", "updates_per_launch", "fetch_bytes_per_launch",
               "write_bytes_per_launch", "traffic_bytes_per_launch", "valu_busy_frac", "valu_issue_frac",
               "valu_utilization", "valu_instr_per_update")


def profile_path(kernel: str, workload: str, store: str) -> str:
    return os.path.join(PROFILES, "%s_%s_%s.md" % (kernel, workload, store))


def profile_figures(kernel: str, workload: str, store: str):
    """The `key: value` figures of a committed profile summary, or None."""
    vals = {}
    try:
        with open(profile_path(kernel, workload, store)) as f:
            for line in f:
                m = re.match(r"^\s*([a-z_]+): (.+?)\s*$", line)
                if m and m.group(1) in FIGURE_KEYS:
                    v = m.group(2)
                    vals[m.group(1)] = v if m.group(1) == "kernel" else float(v)
    except OSError:
        return None
    return vals or None


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
        return dict(parts=a.parts, no_persist=False, reps_per_block=0, block=a.block, kernel=0)
    from mcsolver_amd.autotune import autotune
    r = autotune(b, 1, pt_every, trial_sweeps=64)
    for t in r.trials:
        print("autotune %s: %s" % (t.options, "%.4f ms/sweep" % t.ms_per_sweep if t.ms_per_sweep is not None
                                    else "rejected (%s)" % t.error), file=sys.stderr, flush=True)
    return dict(dict(parts=0, no_persist=False, reps_per_block=0, block=0, kernel=0), **r.best)


def timing_start(sweep: int, every: int) -> int:
    """First sweep of the timing pass: `sweep`, or the next multiple of the
    recompute period when TIMING_SWEEPS from `sweep` would cross one (a
    persistent launch ends at each recompute, and the pass is to be one
    launch).  The Philox counters of skipped sweep indices are never used."""
    if every > 0 and sweep % every + TIMING_SWEEPS > every:
        return (sweep // every + 1) * every
    return sweep


def roofline(b, wl, workload, kernel_for, timing, store):
    """The roofline object of the timing pass (see the module docstring)."""
    from mcsolver_amd import workloads as W
    bpu = W.bytes_per_update(wl.nc, wl.store)
    k_ms, k_launch, k_sweeps, k_upd = timing[0], timing[1], timing[2], timing[3]
    stream = k_launch > k_sweeps  # one launch per colour: the streaming pass
    short = "k_pass" if stream else "k_lds"
    launch_s = k_ms / 1e3 / max(k_launch, 1)
    achieved = (k_upd * bpu) / (k_ms / 1e3) if k_ms > 0 else 0.0
    r = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
         "frac": achieved / HBM_PEAK, "traffic": None, "kernel": kernel_for(stream),
         "avg_launch_us": 1e6 * launch_s, "launches": int(k_launch), "sweeps_per_launch": k_sweeps / max(k_launch, 1),
         "updates_per_launch": k_upd / max(k_launch, 1), "algorithmic_bytes_per_launch": bpu * k_upd / max(k_launch, 1)}
    fig = profile_figures(short, workload, store)
    if fig is not None:
        tr = fig.get("traffic_bytes_per_launch")
        r["traffic"] = tr
        if tr is not None and launch_s > 0:
            r["traffic_frac"] = tr / launch_s / HBM_PEAK
        for k in ("valu_busy_frac", "valu_issue_frac", "valu_utilization", "valu_instr_per_update"):
            if k in fig:
                r[{"valu_busy_frac": "valu_frac"}.get(k, k)] = fig[k]
        r["profile"] = os.path.relpath(profile_path(short, workload, store), ROOT)
        r["profile_kernel"] = fig.get("kernel")
        r["profile_avg_launch_us"] = fig.get("avg_launch_us")
        # what limits the kernel: a measured fraction at 60 % of its peak or
        # more names its unit; below that on both, issue latency (waiting on
        # memory, LDS or barriers with too few waves to cover it)
        vf, hf = r.get("valu_frac"), r.get("traffic_frac")
        if vf is not None and hf is not None:
            r["limiter"] = ("valu" if vf >= hf else "hbm") if max(vf, hf) >= 0.6 else "latency"
    return r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="CPU seconds per point of the CPU leg")
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
    ap.add_argument("--profile-summary", metavar="DIR", default=None,
                    help="print the markdown summary of a rocprofv3 profiling directory and exit")
    a = ap.parse_args()
    if a.profile_summary:
        print(profile_summary(a.profile_summary, a.workload, a.store))
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print("bench.py runs on one GPU only: replicas shard across GPUs with no communication during sweeps, "
              "so N>1 is the same kernel on 1/N of the batch; multi-GPU runs are not part of this build.",
              file=sys.stderr)
        sys.exit(2)
    if a.workload == "cubic_wolff_vs_metropolis":
        wolff_vs_metropolis(a)
        return
    sweep_workload(a)


def sweep_workload(a):
    """K measured sweeps of every replica of the workload (the XY scan swaps
    tempering rungs every sweep), then the roofline timing pass."""
    xy = a.workload == "xy_ladder_scan"
    # CPU leg first: its worker processes fork before the GPU is initialised
    cpu = None
    if not a.no_cpu:
        from oracle import cpu_baseline as cb
        cpu = cb.baseline(a.workload, a.cpu_seconds)

    import torch
    from mcsolver_amd import _lib as L
    from mcsolver_amd import workloads as W

    torch.cuda.set_device(0)
    copy_calibration(a.copy_calib_mib)
    K, Wm, KT = a.steps, a.warmup, TIMING_SWEEPS
    if xy:
        wl = W.build("xy_ladder_scan", nframes=K + Wm + KT, nbins=16)
    else:
        meas = L.MEAS_PAIRS | (L.MEAS_Q if a.workload == "heis_triangular_dmi128" else 0)
        wl = W.build(a.workload, store=a.store, nframes=K + Wm + KT, nbins=16, meas=meas)
    b = wl.batch
    b.init()
    opts = tuned_options(a, b, wl.pt_every)
    b.run(0, Wm, 1, pt_every=wl.pt_every, **opts)
    torch.cuda.synchronize()

    def plan(nsw, sweep0, frame0, timing=None):
        return b.plan(0, nsw, 1, L.METROPOLIS, wl.pt_every, opts["kernel"], opts["block"], opts["reps_per_block"], 0,
                      sweep0, frame0, opts["parts"], opts["no_persist"], timing)

    # timed region: K steps, whole job, no per-launch events.  The warmup ran
    # the same plan shape, so its prepared run (launch plan, tables, device
    # scratch) is reused: the timed call only queues kernel launches.
    st, en = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    p = plan(K, Wm, Wm)
    b.runner(p)
    torch.cuda.synchronize()
    st.record()
    b.run_plan(p)
    en.record()
    torch.cuda.synchronize()
    ms = st.elapsed_time(en)
    ups = wl.updates_per_sweep
    value = ups * K / (ms / 1e3)
    # roofline pass: TIMING_SWEEPS more sweeps with HIP events around every
    # launch of the sweep kernel on the run's stream (the events cost a little
    # wall time, so this pass is kept out of `value`)
    timing = (C.c_double * 4)()
    b.run_plan(plan(KT, timing_start(Wm + K, int(b.b.recompute_every)), Wm + K, timing))
    torch.cuda.synchronize()
    ty = {"f32": "float", "f16": "__half", "i8": "signed char"}[wl.store]
    roof = roofline(b, wl, a.workload,
                    lambda stream: ("k_pass<%d,%s>" if stream else "k_lds<%d,%s>") % (wl.nc, ty), timing, wl.store)
    acc = b.accumulators()
    assert b.error_flag() == 0
    assert (acc[:, L.A_COUNT] == K + Wm + KT).all(), "every temperature row must have one frame per sweep"
    energy_check = _energy_check(b)
    if xy:
        config = {"workload": "XY Binder scan, 7 square lattices x 64-rung PT ladders", "lattices": list(W.XY_SCAN_SIZES),
                  "rungs_per_ladder": 64, "replicas": len(b.reps), "spins_per_sweep": ups, "pt_every": 1,
                  "measure_every": 1, "measured": "E, E^2, |M|, M^2, M^4 per frame (no pair correlations)",
                  "spin_storage": "fp32", "accumulators": "fp64"}
        dtype = "f32"
    else:
        config = {"workload": wl.name + ": " + wl.description, "replicas": len(b.reps), "spins_per_sweep": ups,
                  "measure_every": 1, "measured": "every observable incl. pair correlations" +
                  (" and topological charge" if a.workload == "heis_triangular_dmi128" else ""),
                  "spin_storage": wl.store, "accumulators": "fp64"}
        dtype = {"f32": "f32", "f16": "f32 (f16 storage)", "i8": "f32 (int8 spins)"}[wl.store]
    out = {"metric": METRIC, "value": value, "unit": "spin updates/s", "n_gpus": 1, "steps": K, "warmup": Wm,
           "ms_per_step": ms / K, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": dtype, "data": "synthetic", "config": config,
           "roofline": roof, "cpu_baseline": cpu, "energy_check": energy_check}
    if a.autotune:
        out["plan"] = opts
    if xy and not a.no_plain:
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


# ---------------------------------------------------------------------------
# profile summaries
# ---------------------------------------------------------------------------
def _csv(path):
    with open(path) as f:
        return list(csv.DictReader(f))


def _counters(path):
    """{kernel name: {dispatch id: {counter: value}}} of a counter CSV."""
    d = {}
    for r in _csv(path):
        d.setdefault(r["Kernel_Name"], {}).setdefault(int(r["Dispatch_Id"]), {})[r["Counter_Name"]] = \
            float(r["Counter_Value"])
    return d


def _pick(disp, stream):
    """The profiled launch: the last dispatch of the LDS kernel (the timing
    pass's one launch), or per counter the median over the streaming pass's
    dispatches (all alike: one colour of every replica)."""
    if not stream:
        return disp[max(disp)]
    keys = set().union(*disp.values())
    return {k: float(np.median([v[k] for v in disp.values() if k in v])) for k in keys}


def profile_summary(d: str, workload: str, store: str) -> str:
    ks = _csv(os.path.join(d, "kt", "run_kernel_stats.csv"))
    sweep = [r for r in ks if r["Name"].startswith(("void k_lds", "void k_pass"))]
    kname = max(sweep, key=lambda r: float(r["TotalDurationNs"]))["Name"]
    stream = kname.startswith("void k_pass")
    short = kname.split("(")[0]
    tr = [r for r in _csv(os.path.join(d, "kt", "run_kernel_trace.csv")) if r["Kernel_Name"] == kname]
    durs = {int(r["Dispatch_Id"]): (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3 for r in tr}
    avg_us = durs[max(durs)] if not stream else float(np.median(list(durs.values())))
    bj = json.load(open(os.path.join(d, "bench.json")))
    roof = bj["roofline"]
    upd = roof["updates_per_launch"]
    fig = {"kernel": short, "sweeps_per_launch": roof["sweeps_per_launch"], "avg_launch_us": avg_us,
           "updates_per_launch": upd}
    lines = []
    cal = {}
    for kind, key in (("fetch", "FETCH_SIZE"), ("write", "WRITE_SIZE")):
        c = _counters(os.path.join(d, kind, "run_counter_collection.csv"))
        cps = [v[key] for n, disp in c.items() if "copy" in n.lower() for v in disp.values() if key in v]
        cal[key] = 262144.0 / max(cps)  # the 256 MiB copy: 262,144 KiB each way
        raw = _pick(c[kname], stream)[key]
        fig[kind + "_bytes_per_launch"] = raw * cal[key] * 1024
        lines.append((key, raw, cal[key], max(cps)))
    fig["traffic_bytes_per_launch"] = fig["fetch_bytes_per_launch"] + fig["write_bytes_per_launch"]
    v = _pick(_counters(os.path.join(d, "valu", "run_counter_collection.csv"))[kname], stream)
    cyc = v["GRBM_GUI_ACTIVE"] / NXCD
    fig["valu_busy_frac"] = v["SQ_ACTIVE_INST_VALU"] / NCU / cyc
    fig["valu_issue_frac"] = v["SQ_INSTS_VALU"] / NCU / cyc
    if "SQ_THREAD_CYCLES_VALU" in v:
        fig["valu_utilization"] = v["SQ_THREAD_CYCLES_VALU"] / (v["SQ_ACTIVE_INST_VALU"] * 64)
    fig["valu_instr_per_update"] = v["SQ_INSTS_VALU"] * 64 / upd
    mix = None
    mp_ = os.path.join(d, "mix", "run_counter_collection.csv")
    if os.path.exists(mp_):
        mix = _pick(_counters(mp_)[kname], stream)
    cmds = open(os.path.join(d, "commands.txt")).read().rstrip()
    algo = roof["algorithmic_bytes_per_launch"]
    o = []
    o.append("# %s on %s (%s spins) — rocprofv3 kernel trace and counters\n" % (
        short, workload, bj["config"].get("spin_storage", store)))
    o.append("Workload: %s\n" % bj["config"]["workload"])
    o.append("Commands (run from /tmp with TMPDIR=/tmp; `$R` is the repository root; each counter pass in a run of "
             "its own):\n")
    o.append("\n".join("    " + c for c in cmds.splitlines()) + "\n")
    o.append("Kernel trace (`--kernel-trace --stats`), top kernels of the run:\n")
    o.append("| kernel | calls | average µs | total ms | % |")
    o.append("|---|---|---|---|---|")
    for r in ks[:6]:
        o.append("| %s | %s | %.1f | %.3f | %.1f |" % (r["Name"].split("(")[0], r["Calls"], float(r["AverageNs"]) / 1e3,
                                                   float(r["TotalDurationNs"]) / 1e6, float(r["Percentage"])))
    o.append("")
    o.append("The profiled launch: %s, %.0f site updates per launch, %.1f µs in the kernel trace (bench.py's HIP "
             "events gave %.1f µs in the profiled run).\n" % (
                 "the median over the streaming pass's dispatches (each one colour of every replica)" if stream else
                 "the timing pass's one launch of %d sweeps (the process's last dispatch of the kernel)" %
                 round(roof["sweeps_per_launch"]), upd, avg_us, roof["avg_launch_us"]))
    o.append("HBM traffic (`--pmc FETCH_SIZE`, `--pmc WRITE_SIZE`; the same process copies 256 MiB device to device "
             "as a kernel of known traffic):\n")
    o.append("| counter | raw KiB per launch | copy kernel reads | scale | bytes per launch |")
    o.append("|---|---|---|---|---|")
    for key, raw, sc, cp in lines:
        o.append("| %s | %.1f | %.1f KiB for 262,144 KiB | %.4f | %.4g |" % (key, raw, cp, sc, raw * sc * 1024))
    o.append("")
    o.append("Algorithmic bytes per launch (2 × spin state per site update): %.4g; measured traffic / algorithmic = "
             "%.3f; measured traffic per launch time = %.1f GB/s = %.3f of 8 TB/s.\n" % (
                 algo, fig["traffic_bytes_per_launch"] / algo, fig["traffic_bytes_per_launch"] / avg_us / 1e3,
                 fig["traffic_bytes_per_launch"] / (avg_us * 1e-6) / HBM_PEAK))
    o.append("VALU (`--pmc` %s): cycles = GRBM_GUI_ACTIVE / %d XCDs = %.4g per launch.\n" % (
        " ".join(sorted(v)), NXCD, cyc))
    o.append("| figure | definition | value |")
    o.append("|---|---|---|")
    o.append("| VALU busy | SQ_ACTIVE_INST_VALU / %d CUs / cycles (rocprofv3's VALUBusy for gfx950, as a fraction) | "
             "%.3f |" % (NCU, fig["valu_busy_frac"]))
    o.append("| VALU issue | SQ_INSTS_VALU / %d CUs / cycles (one wave instruction per CU-cycle is the peak without "
             "packed FP32) | %.3f |" % (NCU, fig["valu_issue_frac"]))
    if "valu_utilization" in fig:
        o.append("| VALU lanes active | SQ_THREAD_CYCLES_VALU / (64 × SQ_ACTIVE_INST_VALU) | %.3f |" %
                 fig["valu_utilization"])
    o.append("| VALU instructions per update | 64 × SQ_INSTS_VALU / updates | %.1f |" % fig["valu_instr_per_update"])
    for k in sorted(v):
        if k.startswith("SQ_") and k not in ("SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU", "SQ_THREAD_CYCLES_VALU"):
            o.append("| %s per update | 64 × counter / updates | %.2f |" % (k, v[k] * 64 / upd))
    o.append("")
    if mix is not None:
        o.append("VALU instruction mix (`--pmc` %s), per site update (× 64 lanes / updates):\n" % " ".join(sorted(mix)))
        o.append("| counter | per update |")
        o.append("|---|---|")
        for k in sorted(mix):
            if k != "GRBM_GUI_ACTIVE":
                o.append("| %s | %.2f |" % (k, mix[k] * 64 / upd))
        o.append("")
    o.append("Figures read by bench.py:\n")
    for k in FIGURE_KEYS:
        if k in fig:
            val = fig[k]
            o.append("    %s: %s" % (k, val if isinstance(val, str) else ("%.6g" % val)))
    return "\n".join(o) + "\n"


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
