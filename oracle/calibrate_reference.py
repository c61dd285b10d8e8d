"""Calibration of the CPU baseline against the reference's own engines.

Runs only where oracle/_ref holds the reference's compiled isinglib and
heisenberglib (`make -C oracle ref`, which needs mcsolver's sources).  For
each workload with a reference engine it times, on identical inputs (the
workload's lattice, the same temperatures, sweep counts and glibc seeds),

  reference  mcsolver/isingLib.c or heisenbergLib.c compiled unmodified;
  port       the FP64 restatement (libmcref) that bench.py's CPU leg falls
             back to where oracle/_ref is absent (and always uses for XY);

and prints per-core and aggregate spin updates/s and their ratio, so that a
port figure measured elsewhere converts to the reference's own speed.

    python oracle/calibrate_reference.py [--procs P] [--seconds S]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from oracle import cpu_baseline as cb  # noqa: E402


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--procs", type=int, default=0, help="worker processes (default: usable cores, at most 16)")
    ap.add_argument("--seconds", type=float, default=10.0, help="CPU seconds per point")
    ap.add_argument("--workloads", default="ising_square32,heis_honeycomb64,heis_triangular_dmi128,heis_cubic64")
    a = ap.parse_args()
    rows = []
    for w in a.workloads.split(","):
        if cb.default_kind(w) != "reference":
            print("%s: no reference engine in oracle/_ref, skipped" % w, file=sys.stderr)
            continue
        ref = cb.baseline(w, a.seconds, a.procs or None, "reference")
        port = cb.baseline(w, a.seconds, a.procs or None, "port")
        rows.append({"workload": w, "cores": ref["cores"], "reference": ref["value"], "port": port["value"],
                     "reference_per_core": ref["per_core"], "port_per_core": port["per_core"],
                     "port_over_reference": port["per_core"] / ref["per_core"],
                     "reference_sample": ref["sample"], "port_sample": port["sample"]})
        print(json.dumps(rows[-1]), flush=True)
    print("\nCPU: %s, %d usable cores" % (cpu_model(), len(os.sched_getaffinity(0))))
    print("| workload | cores | reference upd/s | port upd/s | reference per core | port per core | port / reference |")
    print("|---|---|---|---|---|---|---|")
    for r in rows:
        print("| %s | %d | %.3g | %.3g | %.3g | %.3g | %.2f |" % (
            r["workload"], r["cores"], r["reference"], r["port"], r["reference_per_core"], r["port_per_core"],
            r["port_over_reference"]))


if __name__ == "__main__":
    main()
