"""Plan-creation failure diagnostics (FftPlan::c2c_with_diagnostics,
src/plan.rs:227-235; FftPlanCreationError, src/error.rs:187-228): the error
and the blocker FftError::diagnostics gives it (error.rs:498-850), pinned to
the reference's values.  Validation fails before any GPU work, so these run
without a GPU."""
import pytest

import wgpufft_amd as W
from wgpufft_amd import FftConfig, FftDirection, FftPlan, FftPlanCreationError, FftTuning


def fails(make, cfg):
    with pytest.raises(FftPlanCreationError) as e:
        make(cfg)
    return e.value


def test_invalid_axis_is_an_axis_policy_blocker():
    """tests/gpu_c2c.rs:2564-2581."""
    e = fails(FftPlan.c2c_with_diagnostics, FftConfig.new(8).with_axes([1]))
    assert e.kind == "InvalidAxis"
    d = e.diagnostics
    assert d["error"] == "InvalidAxis" and d["route"]["transform"] == "c2c"
    assert any(b["stage"] == "axis-policy" for b in d["blockers"])
    b = d["blockers"][0]
    assert (b["kind"], b["route"], b["layout"], b["required_bytes"], b["actual_bytes"]) == \
        ("Validation", "axis-policy", "axes", 1, 1)
    assert d["active_tuning"]["rader_max_prime"] == 8192


@pytest.mark.parametrize("make,cfg,variant,blocker", [
    (FftPlan.c2c_with_diagnostics, FftConfig.new(0), "ZeroLength",
     ("Validation", "plan-config", "config-shape", "length", 1, 0)),
    (FftPlan.c2c_with_diagnostics, FftConfig.new(8).with_batch(0), "ZeroBatch",
     ("Validation", "plan-config", "config-batch", "batch", 1, 0)),
    (FftPlan.c2c_with_diagnostics, FftConfig.new_nd([4, 4]).with_axes([1, 1]), "DuplicateAxis",
     ("Validation", "axis-policy", "axis-1", "axes", None, None)),
    (FftPlan.c2c_with_diagnostics, FftConfig.new_nd([4, 4]).with_axes([]), "EmptyAxes",
     ("Validation", "axis-policy", "axis-policy", "axes", 1, 0)),
    (FftPlan.c2c_with_diagnostics, FftConfig.new_nd([1, 8]).with_batch(2), "UnsupportedLength",
     ("Unsupported", "mixed-radix", "mixed-radix-factorization", "length", 1, None)),
    (FftPlan.r2c_with_diagnostics, FftConfig.inverse(8), "InvalidRealTransformDirection",
     ("Unsupported", "r2c", "direction", "transform-direction", None, None)),
    (FftPlan.c2r_with_diagnostics, FftConfig.new(8).with_direction(FftDirection.Forward),
     "InvalidRealTransformDirection", ("Unsupported", "c2r", "direction", "transform-direction", None, None)),
    (FftPlan.r2c_with_diagnostics, FftConfig.new_nd([4, 4]).with_axes([1]), "UnsupportedRealAxes",
     ("Unsupported", "real", "real-axis-policy", "real-axes", 2, 1)),
    (FftPlan.c2c_with_diagnostics, FftConfig.new(8).with_tuning(FftTuning().with_force_rader_axes([0])),
     "InvalidTuning", ("Validation", "plan-tuning", "tuning-validation", None, None, None)),
])
def test_creation_blockers(make, cfg, variant, blocker):
    e = fails(make, cfg)
    assert e.kind == variant and e.diagnostics["error"] == variant
    b = e.diagnostics["blockers"][0]
    kind, route, stage, layout, req, act = blocker
    assert (b["kind"], b["route"], b["stage"]) == (kind, route, stage), b
    if layout is not None:
        assert b["layout"] == layout, b
    assert b["required_bytes"] == req and b["actual_bytes"] == act, b
    assert e.diagnostics["stages"] == [] and e.diagnostics["route"]["route"] == route


def test_plain_creation_records_last_error_diagnostics():
    with pytest.raises(W.FftError) as e:
        FftPlan.c2c(FftConfig.new(8).with_axes([3]))
    assert e.value.diagnostics["error"] == "InvalidAxis"
    assert e.value.diagnostics["blockers"][0]["required_bytes"] == 1
    assert e.value.diagnostics["blockers"][0]["actual_bytes"] == 3
