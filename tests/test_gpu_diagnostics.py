"""FftPlan::diagnostics (src/plan.rs:938-1004, src/diagnostics.rs): the JSON
document agrees with the plan's own queries (route, launches, buffer sizes,
tuning) across routes."""
import pytest

from wgpufft_amd import FftConfig, FftPlan, FftPrecision, FftTuning

pytestmark = pytest.mark.gpu

ROUTES = {0: "direct-dft", 1: "mixed-radix", 2: "rader", 3: "bluestein", 4: "axis-sequence"}


@pytest.mark.parametrize("make,kind", [
    (lambda: FftPlan.c2c(FftConfig.new(1024).with_batch(8)), "c2c"),
    (lambda: FftPlan.c2c(FftConfig.new_nd([4096, 2048])), "c2c"),
    (lambda: FftPlan.c2c(FftConfig.new(1009)), "c2c"),
    (lambda: FftPlan.c2c(FftConfig.new(1048583)), "c2c"),
    (lambda: FftPlan.c2c(FftConfig.new(1 << 20)), "c2c"),
    (lambda: FftPlan.c2c(FftConfig.new_nd([16, 16]).with_batch(3)), "c2c"),
    (lambda: FftPlan.c2c(FftConfig.new_nd([256, 256, 256]).with_precision(FftPrecision.F64)), "c2c"),
    (lambda: FftPlan.r2c(FftConfig.new_nd([2048, 2048])), "r2c"),
    (lambda: FftPlan.c2r(FftConfig.inverse_nd([1000, 6])), "c2r"),
])
def test_diagnostics_agree_with_plan(torch_cuda, make, kind):
    plan = make()
    d = plan.diagnostics()
    print(d["route"], [s["kind"] for s in d["stages"]], flush=True)
    assert d["route"]["transform"] == kind
    assert d["route"]["route"] == ROUTES[int(plan.route())]
    assert len(d["stages"]) == plan.launch_count()
    assert [s["label"] for s in d["stages"]] == [plan.launch_describe(i) for i in range(plan.launch_count())]
    req = {b["role"]: b for b in d["buffer_requirements"]}
    assert req["input"]["required_bytes"] == plan.required_input_buffer_size_bytes()
    assert req["output"]["required_bytes"] == plan.required_output_buffer_size_bytes()
    assert req.get("workspace", {"required_bytes": 0})["required_bytes"] == plan.workspace_size_bytes()
    assert d["blockers"] == []
    if kind == "r2c":
        assert req["input"]["format"] == "real-f32" and req["output"]["format"] == "packed-complex-f32"


def test_diagnostics_tuning(torch_cuda):
    t = FftTuning().with_rader_max_prime(101).with_force_bluestein_axes([0]).with_fuse_long_axes(False)
    d = FftPlan.c2c(FftConfig.new(127).with_tuning(t)).diagnostics()
    at = d["active_tuning"]
    assert at["rader_max_prime"] == 101 and at["force_bluestein_axes"] == [0] and at["fuse_long_axes"] is False
    assert d["stages"][0]["kind"] == "bluestein"
