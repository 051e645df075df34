"""Online serving (mifft_server_submit / mifft_server_poll): requests of shapes
never declared beforehand are submitted while earlier ones are in flight;
their plans are built on first sight and kept in a bounded cache.  Every
result matches the FP64 oracle, and a request served amid the mix equals the
same request served alone bit for bit."""
import numpy as np
import pytest

import oracle
from wgpufft_amd import serving as S

pytestmark = pytest.mark.gpu


def host(t, dtype):
    return t.cpu().numpy().view(dtype)


def expected(spec, x_bytes):
    if spec["kind"] == S.R2C:
        return oracle.r2c_nd(x_bytes.view(np.float32).astype(np.float64), spec["shape"], spec["batch"], nthreads=8)
    if spec["precision"] == S.F64:
        return oracle.c2c_nd(x_bytes.view(np.complex128), spec["shape"], batch=spec["batch"], nthreads=8)
    return oracle.c2c_nd(x_bytes.view(np.complex64).astype(np.complex128), spec["shape"], batch=spec["batch"],
                         nthreads=8)


def _f64_fix(specs, bufs):
    import torch
    for s, (x, _) in zip(specs, bufs):
        if s["precision"] == S.F64 and s["kind"] == S.C2C:
            x.view(torch.float64).copy_(torch.rand(x.numel() // 8, dtype=torch.float64, device="cuda") * 2 - 1)


def _check(specs, inputs, outs):
    for k, (s, inp, out) in enumerate(zip(specs, inputs, outs)):
        dt = np.complex128 if (s["precision"] == S.F64 and s["kind"] == S.C2C) else np.complex64
        e = oracle.rel_l2(out.view(dt), expected(s, inp))
        assert e <= (1e-12 if dt == np.complex128 else 1e-6), (k, s, e)


def test_online_submit_poll_matches_oracle_and_alone(torch_cuda):
    import torch
    # the mix includes 2^19..2^20-point 1D requests and 2D up to 1024^2 (run
    # through resident plans beside the batched launches)
    specs = S.request_mix(60, 31)
    specs += [dict(kind=S.C2C, precision=S.F32, shape=[1 << 20], batch=1),
              dict(kind=S.C2C, precision=S.F32, shape=[1024, 1024], batch=1),
              dict(kind=S.R2C, precision=S.F32, shape=[1024, 1024], batch=1)]
    bufs = S.allocate(specs)
    _f64_fix(specs, bufs)
    torch.cuda.synchronize()
    inputs = [host(x, np.uint8).copy() for x, _ in bufs]
    srv = S.FftServer(budget_bytes=16 << 30, streams=4, max_tick_requests=64)
    tickets, done = {}, []
    # first half, then poll once (work in flight), then the rest: shapes of
    # the second half are seen for the first time while the first runs
    half = len(specs) // 2
    for k in range(half):
        tickets[srv.submit(specs[k], bufs[k])] = k
    done += srv.poll()
    q, f, c = srv.pending()
    print(f"after the first half: queued {q}, in flight {f}, completed {c + len(done)}", flush=True)
    for k in range(half, len(specs)):
        tickets[srv.submit(specs[k], bufs[k])] = k
    done += srv.drain()
    assert sorted(d["ticket"] for d in done) == sorted(tickets)
    assert all(d["status"] == 0 and d["done_s"] >= d["admit_s"] >= d["submit_s"] for d in done)
    paths = [d["path"] for d in done]
    assert 0 in paths and 1 in paths
    mixed = [host(y, np.uint8).copy() for _, y in bufs]
    _check(specs, inputs, mixed)
    # each request alone, on a fresh server: the same bits
    for k in range(0, len(specs), 4):
        bufs[k][1].zero_()
        torch.cuda.synchronize()
        alone = S.FftServer(budget_bytes=16 << 30, streams=1, max_tick_requests=8)
        alone.submit(specs[k], bufs[k])
        r = alone.drain()
        assert len(r) == 1 and r[0]["status"] == 0
        assert np.array_equal(host(bufs[k][1], np.uint8), mixed[k]), (k, specs[k])
        del alone
    print(f"online: {len(specs)} requests, {sum(p == 0 for p in paths)} batched, "
          f"{srv.cached_shapes()} cached shapes", flush=True)


def test_online_plan_cache_is_bounded(torch_cuda):
    import torch
    specs = [dict(kind=S.C2C, precision=S.F32, shape=[n], batch=3) for n in (120, 343, 1001, 127, 1009, 4096, 96, 250)]
    bufs = S.allocate(specs)
    torch.cuda.synchronize()
    inputs = [host(x, np.uint8).copy() for x, _ in bufs]
    srv = S.FftServer(budget_bytes=4 << 30, streams=2, max_tick_requests=8)
    srv.set_plan_cache(3)
    for k, s in enumerate(specs):
        srv.submit(s, bufs[k])
        srv.drain()
        assert srv.cached_shapes() <= 3
    _check(specs, inputs, [host(y, np.uint8).copy() for _, y in bufs])


def test_online_rejects_over_budget_at_submission(torch_cuda):
    from wgpufft_amd import FftError
    specs = [dict(kind=S.C2C, precision=S.F32, shape=[65536], batch=8)]
    bufs = S.allocate(specs)
    srv = S.FftServer(budget_bytes=1 << 20, streams=1, max_tick_requests=8)
    with pytest.raises(FftError) as e:
        srv.submit(specs[0], bufs[0])
    assert e.value.kind == "OverBudget"
    assert srv.pending() == (0, 0, 0)
