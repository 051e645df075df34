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
    assert d["stages"][0]["kind"] == "kernel" and d["stages"][0]["algorithm"] == "bluestein"


# Values pinned to the reference's documented diagnostics (tests/gpu_c2c.rs:
# 2525-2545, 2564-2690, 3517-3529; src/diagnostics.rs; src/plan.rs:938-1004):
# route per select_route (c2c.rs:7631-7650), every stage a "kernel", the
# buffer roles and formats, no blockers for a plan alone.
@pytest.mark.parametrize("make,route,roles", [
    (lambda: FftPlan.c2c(FftConfig.new(1024).with_batch(8)), "mixed-radix",
     {"input": "complex-f32", "output": "complex-f32", "helper:twiddle-luts-total": "complex-f32"}),
    (lambda: FftPlan.c2c(FftConfig.new_nd([4096, 2048])), "mixed-radix",
     {"input": "complex-f32", "output": "complex-f32", "workspace": "complex-f32",
      "helper:twiddle-luts-total": "complex-f32"}),
    (lambda: FftPlan.c2c(FftConfig.new(1009).with_batch(2)), "rader",
     {"input": "complex-f32", "output": "complex-f32", "helper:twiddle-luts-total": "complex-f32"}),
    (lambda: FftPlan.c2c(FftConfig.new(1009).with_tuning(FftTuning().with_force_bluestein_axes([0]))), "bluestein",
     {"input": "complex-f32", "output": "complex-f32", "helper:twiddle-luts-total": "complex-f32"}),
    (lambda: FftPlan.c2c(FftConfig.new_nd([1009, 8])), "axis-sequence",
     {"input": "complex-f32", "output": "complex-f32"}),
    (lambda: FftPlan.c2c(FftConfig.new(1)), "direct-dft", {"input": "complex-f32", "output": "complex-f32"}),
    (lambda: FftPlan.c2c(FftConfig.new_nd([256, 256, 256]).with_precision(FftPrecision.F64)), "mixed-radix",
     {"input": "complex-f64", "output": "complex-f64", "workspace": "complex-f64"}),
    (lambda: FftPlan.r2c(FftConfig.new_nd([2048, 2048])), "mixed-radix",
     {"input": "real-f32", "output": "packed-complex-f32", "workspace": "complex-f32"}),
])
def test_diagnostics_pinned_values(torch_cuda, make, route, roles):
    plan = make()
    d = plan.diagnostics()
    assert d["route"]["route"] == route
    assert d["route"]["large_route_mode"] == "normal" and d["route"]["execution_kind"] == "normal"
    assert d["stages"] and all(s["kind"] == "kernel" and s["route"] == route for s in d["stages"])
    req = {b["role"]: b["format"] for b in d["buffer_requirements"]}
    for role, fmt in roles.items():
        assert req.get(role) == fmt, (role, req)
    assert len(d["buffer_requirements"]) >= 2 and d["blockers"] == []


def test_execution_failure_diagnostics_pinned(torch_cuda):
    """tests/gpu_c2c.rs:2583-2690: an input view 8 bytes short fails with
    BufferViewTooSmall and one Validation blocker (route mixed-radix, stage
    input-logical-view, layout contiguous, required and actual bytes)."""
    import torch
    from wgpufft_amd import FftError
    plan = FftPlan.c2c(FftConfig.new(8))
    required = plan.required_buffer_size_bytes()
    small = torch.zeros(required - 8, dtype=torch.uint8, device="cuda")
    ok = torch.zeros(required, dtype=torch.uint8, device="cuda")
    with pytest.raises(FftError) as e:
        plan.execute(small, ok)
    assert e.value.kind == "BufferViewTooSmall"
    bl = [b for b in e.value.diagnostics["blockers"]
          if b["kind"] == "Validation" and b["route"] == "mixed-radix" and b["stage"] == "input-logical-view"
          and b["layout"] == "contiguous" and b["required_bytes"] == required and b["actual_bytes"] == required - 8]
    assert len(bl) == 1, e.value.diagnostics["blockers"]
    # the output side the same way
    with pytest.raises(FftError) as e:
        plan.execute(ok, small)
    assert any(b["stage"] == "output-logical-view" and b["actual_bytes"] == required - 8
               for b in e.value.diagnostics["blockers"])


def test_diagnostics_for_workspace_and_views(torch_cuda):
    """FftPlan::diagnostics_for_workspace / _for_views (plan.rs:1023-1300)."""
    import torch
    from wgpufft_amd import FftLogicalLayout
    plan = FftPlan.c2c(FftConfig.new_nd([1024, 512]))
    need = plan.required_buffer_size_bytes()
    ws_need = plan.workspace_size_bytes()
    assert ws_need > 0
    buf = torch.zeros(need, dtype=torch.uint8, device="cuda")
    ok = plan.diagnostics_for_workspace(buf, buf, (buf.data_ptr(), ws_need))
    assert ok["blockers"] == []
    d = plan.diagnostics_for_workspace(buf, buf, (buf.data_ptr(), ws_need - 16))
    ws = [b for b in d["blockers"] if b["kind"] == "Workspace"]
    assert len(ws) == 1 and ws[0]["stage"] == "workspace" and ws[0]["helper_buffer"] == "workspace"
    assert ws[0]["required_bytes"] == ws_need and ws[0]["actual_bytes"] == ws_need - 16
    d = plan.diagnostics_for_workspace((buf.data_ptr() + 4, need), buf, (buf.data_ptr(), ws_need))
    assert [b["kind"] for b in d["blockers"]] == ["Alignment"]
    # logical layouts
    line = FftPlan.c2c(FftConfig.new(64).with_batch(4))
    big = torch.zeros(64 * 4 * 8 * 3, dtype=torch.uint8, device="cuda")
    assert line.diagnostics_for_views(big, FftLogicalLayout(0, 2), big, None)["blockers"] == []
    d = line.diagnostics_for_views(big, FftLogicalLayout(0, 0), big, None)
    assert [(b["kind"], b["layout"]) for b in d["blockers"]] == [("Layout", "element-stride")]
    d = line.diagnostics_for_views(big, FftLogicalLayout(0, 1, 10), big, None)
    b = d["blockers"][0]
    assert (b["kind"], b["layout"], b["required_bytes"], b["actual_bytes"]) == ("Layout", "batch-stride", 64, 10)
    d = line.diagnostics_for_views(big, FftLogicalLayout(0, 4), big, None)
    b = [x for x in d["blockers"] if x["stage"] == "input-logical-view"]
    assert b and b[0]["layout"] == "strided" and b[0]["required_bytes"] == (63 * 4 * 4 + 4) * 8
    d = line.diagnostics_for_views(big, FftLogicalLayout(0, 2), big, None, workspace=16)
    assert [x["kind"] for x in d["blockers"]] == ["Workspace"]
