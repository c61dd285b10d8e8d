"""Host-side checks of the plan autotuner (no GPU needed)."""
import pytest

from mcsolver_amd.autotune import DEFAULT_CANDIDATES, autotune


def test_default_plan_is_a_candidate():
    # the tuner can always fall back to what the planner would pick by itself
    assert dict(parts=0, no_persist=False, reps_per_block=0) in list(DEFAULT_CANDIDATES)
    assert len({tuple(sorted(c.items())) for c in DEFAULT_CANDIDATES}) == len(DEFAULT_CANDIDATES)


@pytest.mark.parametrize("kw", [dict(trial_sweeps=0), dict(sweeps_per_frame=0), dict(candidates=[])])
def test_bad_arguments_rejected_before_any_run(kw):
    with pytest.raises(ValueError):
        autotune(None, **kw)
