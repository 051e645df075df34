"""mifft_execute_static_scratch, called through the C ABI: the same plans run
with the library's static device scratch as workspace give the same bytes as
mifft_execute_with_workspace with a caller workspace, on two calls in a row,
the second with a smaller input; an execution needing more scratch than the
static array holds is refused before any GPU work."""
import ctypes

import numpy as np
import pytest

import oracle
from tests.helpers import empty_dev, from_dev, seeded_complex, to_dev
import wgpufft_amd as W
from wgpufft_amd import FftConfig, FftPlan, FftTuning, _lib

pytestmark = pytest.mark.gpu

SCRATCH = 80 << 20


def _ws_run(plan, x_dev, out_dev):
    lib = _lib.load()
    ws = empty_dev(max(16, plan.workspace_size_bytes()))
    st = W._stream(None)
    rc = lib.mifft_execute_with_workspace(plan._h, st, x_dev.data_ptr(), x_dev.numel(), out_dev.data_ptr(),
                                          out_dev.numel(), ws.data_ptr(), ws.numel())
    assert rc == 0, W.FftError(rc)


def _scratch_run(plan, x_dev, out_dev):
    lib = _lib.load()
    st = W._stream(None)
    return lib.mifft_execute_static_scratch(plan._h, st, x_dev.data_ptr(), x_dev.numel(), out_dev.data_ptr(),
                                            out_dev.numel())


def _case(kind, cfg, seed, in_place):
    import torch
    plan = {"c2c": FftPlan.c2c, "r2c": FftPlan.r2c}[kind](cfg)
    n_in = plan.required_input_buffer_size_bytes()
    n_out = plan.required_output_buffer_size_bytes()
    if kind == "r2c":
        x = oracle.normal_f32(seed, n_in // 4)
    else:
        x = seeded_complex(seed, n_in // 8)
    size = max(n_in, n_out)
    a_in = empty_dev(size)
    a_in[:n_in].copy_(to_dev(x))
    b_in = a_in.clone()
    a_out = a_in if in_place else empty_dev(n_out)
    b_out = b_in if in_place else empty_dev(n_out)
    _ws_run(plan, a_in, a_out)
    rc = _scratch_run(plan, b_in, b_out)
    torch.cuda.synchronize()
    assert rc == 0, W.FftError(rc)
    want = from_dev(a_out[:n_out], np.uint8)
    got = from_dev(b_out[:n_out], np.uint8)
    assert np.array_equal(got, want), (kind, cfg.shape(), plan.describe())
    # and the values are the transform (oracle, relative L2 <= 1e-6)
    if kind == "c2c":
        y = got.view(np.complex64)
        ref = oracle.c2c_nd(x.astype(np.complex128), cfg.shape(), batch=cfg.batch(), nthreads=8)
    else:
        y = got.view(np.complex64)
        ref = oracle.r2c_nd(x, cfg.shape(), batch=cfg.batch(), nthreads=8)
    e = oracle.rel_l2(y, ref)
    print(f"static scratch {kind} {cfg.shape()} b={cfg.batch()} in_place={in_place}: bit-equal, rel_l2={e:.3e} "
          f"[{plan.describe()}]", flush=True)
    assert e <= 1e-6
    return plan.workspace_size_bytes()


def test_static_scratch_matches_workspace_execution(torch_cuda):
    # the first call is the largest the scratch takes (the 4096 x 2048 pair's
    # padded 64 MiB ping-pong workspace); the second a smaller input
    ws1 = _case("c2c", FftConfig.new_nd([4096, 2048]), 11, True)
    assert 0 < ws1 <= SCRATCH
    ws2 = _case("c2c", FftConfig.new_nd([1024, 512]), 12, True)
    assert 0 < ws2 < ws1
    # a multi-pass Bluestein axis, a three-pass split, a real transform in
    # place (input copy in the scratch) and out of place
    _case("c2c", FftConfig.new(10007).with_batch(3).with_tuning(FftTuning().with_force_bluestein_axes([0])),
          13, False)
    _case("c2c", FftConfig.new(1 << 20), 14, True)
    _case("r2c", FftConfig.new_nd([2048, 2048]), 15, True)
    _case("r2c", FftConfig.new(1 << 20).with_batch(2), 16, False)


def test_static_scratch_refuses_what_does_not_fit(torch_cuda):
    import torch
    plan = FftPlan.c2c(FftConfig.new_nd([4096, 4096]))  # 128 MiB workspace
    assert plan.workspace_size_bytes() > SCRATCH
    buf = empty_dev(plan.required_buffer_size_bytes())
    rc = _scratch_run(plan, buf, buf)
    torch.cuda.synchronize()
    assert W.FftError(rc).kind == "WorkspaceTooSmall"
    detail = _lib.load().mifft_last_error_detail().decode()
    assert "static scratch holds" in detail, detail
