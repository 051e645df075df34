"""GPU tests of the continuous-batching server (mifft_server_*).

A seeded trace of mixed requests (1D f32/f64 C2C with smooth, prime and
Bluestein lengths, 2D f32 C2C, R2C) is replayed with Poisson arrivals; every
result must match the FP64 oracle, and must be bit-identical to the same
request served alone (the server's output does not depend on what else is in
flight)."""
import numpy as np
import pytest

import oracle
from wgpufft_amd import serving as S

pytestmark = pytest.mark.gpu


def host(t, dtype):
    return t.cpu().numpy().view(dtype)


def expected(spec, x_bytes):
    if spec["kind"] == S.R2C:
        x = x_bytes.view(np.float32).astype(np.float64)
        return oracle.r2c_nd(x, spec["shape"], spec["batch"], nthreads=8)
    if spec["precision"] == S.F64:
        xc = x_bytes.view(np.complex128)
    else:
        xc = x_bytes.view(np.complex64).astype(np.complex128)
    return oracle.c2c_nd(xc, spec["shape"], batch=spec["batch"], nthreads=8)


def small_mix(count, seed, large=False):
    specs = [s for s in S.request_mix(count * 3, seed) if np.prod(s["shape"]) <= 300_000][:count]
    if large:
        # the serving mix's largest requests: 2^19..2^20-point 1D and 1024^2 2D
        specs += [dict(kind=S.C2C, precision=S.F32, shape=[1 << 19], batch=1),
                  dict(kind=S.C2C, precision=S.F32, shape=[1 << 20], batch=1),
                  dict(kind=S.C2C, precision=S.F64, shape=[786432], batch=1),
                  dict(kind=S.C2C, precision=S.F32, shape=[1024, 1024], batch=1),
                  dict(kind=S.R2C, precision=S.F32, shape=[1024, 1024], batch=1)]
    return specs


def test_serve_mix_matches_oracle_and_alone(torch_cuda):
    import torch
    specs = small_mix(40, 12, large=True)
    bufs = S.allocate(specs)
    # f64 inputs: the f32 fill gives raw bytes; replace by finite f64 values
    for s, (x, _) in zip(specs, bufs):
        if s["precision"] == S.F64 and s["kind"] == S.C2C:
            n = x.numel() // 8
            x.view(torch.float64).copy_(torch.rand(n, dtype=torch.float64, device="cuda") * 2 - 1)
    torch.cuda.synchronize()
    inputs = [host(x, np.uint8).copy() for x, _ in bufs]
    arr = S.poisson_arrivals(len(specs), 20_000.0, seed=3)
    srv = S.FftServer(budget_bytes=8 << 30, streams=4, max_tick_requests=64)
    res = srv.replay(specs, bufs, arr, mode=0)
    assert (res["status"] == 0).all()
    summ = S.latency_summary(res)
    print("serve mix", summ, flush=True)
    assert summ["batched_fraction"] > 0.0
    mixed = [host(y, np.uint8).copy() for _, y in bufs]
    for k, (s, inp, out) in enumerate(zip(specs, inputs, mixed)):
        ref = expected(s, inp)
        dt = np.complex128 if (s["precision"] == S.F64 and s["kind"] == S.C2C) else np.complex64
        got = out.view(dt)
        tol = 1e-12 if dt == np.complex128 else 1e-6
        e = oracle.rel_l2(got, ref)
        assert e <= tol, (k, s, e)
    # alone: each request replayed by itself gives the same bits
    for k in range(0, len(specs), 3):
        bufs[k][1].zero_()
        torch.cuda.synchronize()
        r1 = srv.replay([specs[k]], [bufs[k]], [0.0], mode=0)
        assert r1["status"][0] == 0
        assert np.array_equal(host(bufs[k][1], np.uint8), mixed[k]), (k, specs[k])


def test_serve_sequential_baseline_matches_oracle(torch_cuda):
    import torch
    specs = [s for s in small_mix(20, 21) if s["precision"] == S.F32]
    bufs = S.allocate(specs)
    torch.cuda.synchronize()
    inputs = [host(x, np.uint8).copy() for x, _ in bufs]
    srv = S.FftServer(budget_bytes=8 << 30, streams=2, max_tick_requests=16)
    res = srv.replay(specs, bufs, S.poisson_arrivals(len(specs), 20_000.0), mode=1)
    assert (res["status"] == 0).all() and (res["path"] == 1).all()
    for s, inp, (_, y) in zip(specs, inputs, bufs):
        assert oracle.rel_l2(host(y, np.complex64), expected(s, inp)) <= 1e-6


def test_serve_rejects_request_over_budget(torch_cuda):
    specs = [dict(kind=S.C2C, precision=S.F32, shape=[65536], batch=8),
             dict(kind=S.C2C, precision=S.F32, shape=[1024], batch=1)]
    bufs = S.allocate(specs)
    srv = S.FftServer(budget_bytes=1 << 20, streams=1, max_tick_requests=8)
    res = srv.replay(specs, bufs, [0.0, 0.0], mode=0)
    assert res["status"][0] != 0 and res["path"][0] == 2
    assert res["status"][1] == 0


def test_serve_batch_limit_routes_large_requests(torch_cuda):
    """Requests at or above the batch limit run through their resident plans
    beside the batched launch; every result still matches the oracle."""
    import torch
    specs = [s for s in small_mix(30, 5) if s["precision"] == S.F32]
    bufs = S.allocate(specs)
    torch.cuda.synchronize()
    inputs = [host(x, np.uint8).copy() for x, _ in bufs]
    limit = 4096
    srv = S.FftServer(budget_bytes=8 << 30, streams=4, max_tick_requests=32, batch_max_points=limit)
    res = srv.replay(specs, bufs, S.poisson_arrivals(len(specs), 20_000.0, seed=9), mode=0)
    assert (res["status"] == 0).all()
    points = np.array([int(np.prod(s["shape"])) * s["batch"] for s in specs])
    assert (res["path"][points >= limit] == 1).all()
    assert (res["path"][points < limit] == 0).any() and (points >= limit).any()
    for s, inp, (_, y) in zip(specs, inputs, bufs):
        assert oracle.rel_l2(host(y, np.complex64), expected(s, inp)) <= 1e-6, s
