"""Measured choice of the launch plan of a batch.

mcMain.py runs one process per temperature (win.py:56-156 hands (T, H) points
to a multiprocessing Pool) and has nothing to tune.  Here one call runs every
replica, and the planner of libmcgpu deals replicas to workgroups by a fixed
rule (sites per CU, automatic splits of large replicas, persistent launches
where every workgroup is resident).  autotune() replaces that rule by a
measurement: it times a few trial sweeps of each candidate plan on the batch
itself and returns the fastest.  The plan dimensions (Batch.run options):

  parts           workgroups per LDS replica (the tile: 0 the planner's
                  rule, 1 whole replicas, 2 or 4 parts with halo groups)
  no_persist      persistent launches (barrier and tempering inside) or not
  reps_per_block  the planner's packing of replicas per workgroup, or one
  block           threads per workgroup of the LDS kernel (0: the planner's
                  1024, or 256 / 512)
  kernel          the LDS-resident kernel or the streaming colour pass

autotune() searches them one dimension at a time from the planner's plan
(coordinate descent, each dimension once in the order above), or times an
explicit candidate list.  The colour schedule is the graph's colouring
(engine.compile_graph); any proper colouring samples the same distribution,
but it is part of the trajectory, so it is not a plan option.

The choice does not change the trajectory.  Philox counters are keyed by
replica, site group, sweep and colour, never by workgroup or thread, so every
plan gives the same spins (the GPU tests compare plans bit for bit).  The
tracked energy and magnetisation are FP64 sums of per-thread FP32 partial
sums, and which sites share a thread depends on the plan, so frame
observables agree to FP32 rounding (about 1e-7 relative) and a tempering swap
could differ only where its acceptance ratio lies within that rounding of its
uniform draw.  The trials do advance the batch, so every device buffer a run
writes is saved first and restored after each trial.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import torch

from . import _lib as L
from .engine import Batch

# (parts, no_persist, reps_per_block): automatic splits or none or two parts,
# persistent launches on or off, the planner's packing or one replica per
# workgroup (an explicit grid for autotune(candidates=...))
DEFAULT_CANDIDATES: Sequence[Dict] = tuple(
    dict(parts=p, no_persist=np_, reps_per_block=rb)
    for p in (0, 1, 2) for np_ in (False, True) for rb in (0, 1))

# the dimensions of the coordinate search, each with its values (the first
# is the planner's default)
DIMENSIONS: Sequence = (
    ("parts", (0, 1, 2, 4)),
    ("no_persist", (False, True)),
    ("reps_per_block", (0, 1)),
    ("block", (0, 256, 512)),
    ("kernel", (L.KERNEL_AUTO, L.KERNEL_STREAM)),
)


@dataclass
class Trial:
    options: Dict
    ms_per_sweep: Optional[float]      # None: the plan was rejected
    error: str = ""


@dataclass
class TuneResult:
    best: Dict
    trials: List[Trial] = field(default_factory=list)


def _mutable(b: Batch) -> List[torch.Tensor]:
    """Every device buffer that mcg_run writes."""
    t = [b.spins, b.state, b.acc, b.bins, b.gacc, b.pt_acc, b.err, b.rung, b.who]
    return t + [x for x in (b.wstamp, b.wlist) if x is not None]


def autotune(b: Batch, sweeps_per_frame: int = 1, pt_every: int = 0, trial_sweeps: int = 64,
             candidates: Optional[Sequence[Dict]] = None) -> TuneResult:
    """Time Batch.run options on trial_sweeps measured sweeps (frames every
    sweeps_per_frame sweeps, tempering every pt_every sweeps, as the
    production run will have them) and return the fastest: the coordinate
    search over DIMENSIONS, or every plan of `candidates`.  Each plan runs
    once untimed, then once timed with HIP events on the current stream.
    A candidate the planner rejects, or whose run raises the device error flag,
    is skipped.  The batch's spins, tracked state, accumulators, bins, ladders
    and error flag are as before the call when it returns."""
    if trial_sweeps < 1 or sweeps_per_frame < 1:
        raise ValueError("trial_sweeps and sweeps_per_frame must be >= 1")
    if candidates is not None and not list(candidates):
        raise ValueError("no candidate plans")
    nframe = max(1, trial_sweeps // sweeps_per_frame)
    bufs = _mutable(b)
    saved = [x.clone() for x in bufs]
    nframes0 = b.b.nframes
    sweep0, cl_end0 = b._next_sweep, b._wolff_cl_end
    b.set_nframes(nframe)
    s = torch.cuda.current_stream()
    out = TuneResult(best={})
    seen = {}

    def trial(opt):
        """ms per sweep of plan opt (None: rejected), each plan timed once."""
        key = tuple(sorted(opt.items()))
        if key in seen:
            return seen[key]
        ms = None
        try:
            times = []
            for _ in range(2):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(s)
                b.run(0, nframe, sweeps_per_frame, pt_every=pt_every, sweep0=sweep0, **opt)
                e1.record(s)
                e1.synchronize()
                times.append(e0.elapsed_time(e1))
                flag = b.error_flag()
                for x, y in zip(bufs, saved):
                    x.copy_(y)
                if flag:
                    raise RuntimeError("device error flag 0x%x" % flag)
            ms = times[-1] / (nframe * sweeps_per_frame)
            out.trials.append(Trial(dict(opt), ms))
        except RuntimeError as e:
            for x, y in zip(bufs, saved):
                x.copy_(y)
            out.trials.append(Trial(dict(opt), None, str(e)))
        seen[key] = ms
        return ms

    try:
        if candidates is not None:
            for opt in candidates:
                trial(dict(opt))
        else:
            # coordinate descent from the planner's plan
            cur = {k: vals[0] for k, vals in DIMENSIONS}
            best_ms = trial(cur)
            for k, vals in DIMENSIONS:
                for v in vals[1:]:
                    opt = dict(cur, **{k: v})
                    ms = trial(opt)
                    if ms is not None and (best_ms is None or ms < best_ms):
                        cur, best_ms = opt, ms
    finally:
        for x, y in zip(bufs, saved):
            x.copy_(y)
        b.set_nframes(nframes0)
        b._next_sweep, b._wolff_cl_end = sweep0, cl_end0
        torch.cuda.synchronize()
    ok = [t for t in out.trials if t.ms_per_sweep is not None]
    if not ok:
        raise RuntimeError("autotune: every candidate plan failed (%s)" %
                           "; ".join(t.error for t in out.trials))
    out.best = dict(min(ok, key=lambda t: t.ms_per_sweep).options)
    return out
