"""Plans beyond the u32 index space: the large-chunk route (batch chunks of at
most 2^31 elements run at 64-bit byte offsets; the reference's LargeChunkPlan,
src/runtime/large_chunk.rs:5-120, large_policy.rs:325).  Full-size buffers of
more than 2^32 elements on the device, sampled lines against the FP64 oracle
(relative L2 <= 1e-6), lines on both sides of every chunk boundary and the
last line, and a sample equal bit for bit to a standalone plan of the same
lines."""
import numpy as np
import pytest

import oracle
from tests.helpers import empty_dev, from_dev, to_dev
import wgpufft_amd as W
from wgpufft_amd import FftConfig, FftPlan

pytestmark = pytest.mark.gpu

CHUNK = 1 << 31  # elements per chunk


def test_batched_1d_beyond_u32(torch_cuda):
    import torch
    n = 1024
    batch = (1 << 22) + 5  # 4,294,972,416 elements (32 GiB), chunks of 2^21 lines
    total = n * batch
    assert total > 0xFFFFFFFF
    plan = FftPlan.c2c(FftConfig.new(n).with_batch(batch))
    d = plan.diagnostics()
    assert d["route"]["large_route_mode"] == "large-chunk", d["route"]
    assert plan.describe().count("large-chunk") == 3, plan.describe()
    assert plan.required_buffer_size_bytes() == 8 * total
    buf = torch.empty(8 * total, dtype=torch.uint8, device="cuda")
    W.fill_uniform_f32(buf, 77)
    per_chunk = CHUNK // n
    rng = np.random.default_rng(5)
    lines = sorted(set([0, 1, per_chunk - 1, per_chunk, per_chunk + 1, 2 * per_chunk - 1, 2 * per_chunk,
                        batch - 2, batch - 1] + [int(v) for v in rng.integers(0, batch, 24)]))
    lb = 8 * n
    x = {ln: from_dev(buf[ln * lb:(ln + 1) * lb], np.complex64).copy() for ln in lines}
    # 8 lines across the first chunk boundary, for the standalone comparison
    s0 = per_chunk - 4
    span = buf[s0 * lb:(s0 + 8) * lb].clone()
    torch.cuda.synchronize()
    plan.execute_in_place(buf)
    torch.cuda.synchronize()
    worst = 0.0
    for ln in lines:
        y = from_dev(buf[ln * lb:(ln + 1) * lb], np.complex64)
        e = oracle.rel_l2(y, oracle.c2c_nd(x[ln].astype(np.complex128), [n]))
        worst = max(worst, e)
        assert e <= 1e-6, (ln, e)
    small = FftPlan.c2c(FftConfig.new(n).with_batch(8))
    out = empty_dev(span.numel())
    small.execute(span, out)
    torch.cuda.synchronize()
    assert np.array_equal(from_dev(out, np.uint8), from_dev(buf[s0 * lb:(s0 + 8) * lb], np.uint8))
    print(f"[1024] x {batch}: {len(lines)} sampled lines, worst rel_l2={worst:.3e}; "
          f"chunk boundary equals a standalone plan bit for bit [{plan.describe()}]", flush=True)
    del buf


def test_r2c_2d_batches_beyond_u32(torch_cuda):
    import torch
    shape = [2048, 1024]
    batch = 2049  # 4,297,064,448 real elements; chunks of 1024 batches
    per = shape[0] * shape[1]
    total = per * batch
    assert total > 0xFFFFFFFF
    r2c = FftPlan.r2c(FftConfig.new_nd(shape).with_batch(batch))
    assert r2c.diagnostics()["route"]["large_route_mode"] == "large-chunk"
    pin = 4 * per
    pout = 8 * (shape[0] // 2 + 1) * shape[1]
    assert r2c.required_input_buffer_size_bytes() == pin * batch
    assert r2c.required_output_buffer_size_bytes() == pout * batch
    src = torch.empty(pin * batch, dtype=torch.uint8, device="cuda")
    dst = torch.empty(pout * batch, dtype=torch.uint8, device="cuda")
    W.fill_normal_f32(src, 78)
    picks = [0, 1023, 1024, 2047, 2048]
    xs = {b: from_dev(src[b * pin:(b + 1) * pin], np.float32).copy() for b in picks}
    torch.cuda.synchronize()
    r2c.execute(src, dst)
    torch.cuda.synchronize()
    for b in picks:
        y = from_dev(dst[b * pout:(b + 1) * pout], np.complex64)
        e = oracle.rel_l2(y, oracle.r2c_nd(xs[b], shape, nthreads=16))
        print(f"R2C {shape} batch {b} of {batch}: rel_l2={e:.3e}", flush=True)
        assert e <= 1e-6, (b, e)
    del src, dst


def test_strided_views_of_chunked_plans_are_refused(torch_cuda):
    import torch
    from wgpufft_amd import FftLogicalLayout
    plan = FftPlan.c2c(FftConfig.new(1 << 20).with_batch(1 << 12))
    assert plan.diagnostics()["route"]["large_route_mode"] == "large-chunk"
    tiny = torch.empty(16, dtype=torch.uint8, device="cuda")
    with pytest.raises(W.FftError) as e:
        plan.execute_logical_views((tiny.data_ptr(), 1 << 40), FftLogicalLayout(0, 2),
                                   (tiny.data_ptr(), 1 << 40), None)
    assert e.value.kind == "BufferLayoutTooLarge"
