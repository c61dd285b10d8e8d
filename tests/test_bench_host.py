"""bench.py's host logic (no GPU): the committed profile summaries it reads
for the roofline fields, and the placement of the roofline timing pass."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402

# every (kernel, workload, storage) a default bench line of each workload looks up
LOOKUPS = [("k_lds", "xy_ladder_scan", "f32"), ("k_lds", "ising_square32", "i8"),
           ("k_lds", "heis_honeycomb64", "f32"), ("k_pass", "heis_triangular_dmi128", "f32"),
           ("k_lds", "heis_triangular_dmi128", "f16"), ("k_pass", "heis_cubic64", "f32")]


@pytest.mark.parametrize("kernel,workload,store", LOOKUPS)
def test_profile_summaries_parse(kernel, workload, store):
    f = bench.profile_figures(kernel, workload, store)
    assert f is not None, bench.profile_path(kernel, workload, store)
    assert f["kernel"].startswith("void " + kernel + "<")
    for k in ("avg_launch_us", "updates_per_launch", "fetch_bytes_per_launch", "write_bytes_per_launch",
              "traffic_bytes_per_launch", "valu_busy_frac", "valu_issue_frac", "valu_instr_per_update"):
        assert f[k] > 0, k
    assert abs(f["traffic_bytes_per_launch"] - f["fetch_bytes_per_launch"] - f["write_bytes_per_launch"]) < \
        1e-4 * f["traffic_bytes_per_launch"]
    assert 0 < f["valu_busy_frac"] <= 1.0 and 0 < f["valu_issue_frac"] <= 1.0
    # the LDS kernels are profiled on the timing pass's one launch
    if kernel == "k_lds":
        assert f["sweeps_per_launch"] == bench.TIMING_SWEEPS
    # every summary names the commands that made it, and only committed files
    text = open(bench.profile_path(kernel, workload, store)).read()
    assert "rocprofv3 --kernel-trace --stats" in text and "--pmc FETCH_SIZE" in text and "--pmc WRITE_SIZE" in text
    assert "_work" not in text


def test_timing_pass_is_one_launch():
    every = 1000
    for s in (0, 25, 600, 601, 999, 1000, 2200, 2700):
        t = bench.timing_start(s, every)
        assert t >= s and t - s < every
        assert t % every + bench.TIMING_SWEEPS <= every  # no recompute boundary inside
    assert bench.timing_start(601, every) == 1000
    assert bench.timing_start(123, 0) == 123


def test_roofline_fields_from_profile():
    class W:
        nc, store = 2, "f32"
    t = [1000.0 * 6.4127, 1.0, 400.0, 866713600.0]  # ms, launches, sweeps, updates of the timing pass
    r = bench.roofline(None, W, "xy_ladder_scan", lambda s: "k_lds<2,float>", t, "f32")
    assert r["bound"] == "hbm" and r["unit"] == "GB/s"
    assert r["achieved"] == pytest.approx(866713600.0 * 16 / 6.4127 / 1e9, rel=1e-9)
    assert r["frac"] == pytest.approx(r["achieved"] / 8000.0)
    assert r["traffic"] > 0 and 0 < r["traffic_frac"] < 1
    assert 0.5 < r["valu_frac"] < 1 and r["limiter"] in ("valu", "latency")
    assert r["profile"] == os.path.join("profiles", "k_lds_xy_ladder_scan_f32.md")
