"""GPU tests of the persistent batched launch (FftBatch / mifft_batch_*).

For these lengths the batched planner picks the same kernels as a standalone
plan, so every transform must be bit-identical to its standalone plan (the
batched kernel runs the same tile functions on the same tables); its accuracy
against the FP64 oracle is checked as well.
The lengths are those of the uneven-batch workload: 120, 343, 1001
(smooth), 127 (Rader), 1009 (forced Bluestein), 4096 and 65536 (four-step).
"""
import numpy as np
import pytest

import oracle
from tests.helpers import c2c_expected, empty_dev, from_dev, seeded_complex, to_dev
from wgpufft_amd import FftBatch, FftConfig, FftDirection, FftError, FftPlan, FftTuning, Normalization

pytestmark = pytest.mark.gpu

LENGTHS = [120, 343, 1001, 127, 1009, 4096, 65536]


def make_items(seed, count, max_batch, inverse_mix=False):
    rng = np.random.default_rng(seed)
    items = []
    for t in range(count):
        n = int(LENGTHS[rng.integers(len(LENGTHS))])
        b = int(rng.integers(1, max_batch + 1))
        if n == 65536:
            b = min(b, 3)
        x = seeded_complex(10_000 + 7 * t, n * b)
        d_in = to_dev(x)
        in_place = bool(rng.integers(2))
        d_out = d_in if in_place else empty_dev(x.nbytes)
        direction = FftDirection.Inverse if (inverse_mix and rng.integers(2)) else FftDirection.Forward
        norm = Normalization(int(rng.integers(4))) if inverse_mix else Normalization.Inverse
        items.append(dict(n=n, batch=b, x=x, input=d_in, output=d_out, direction=direction,
                          normalization=norm, algorithm=2 if n == 1009 else 0, in_place=in_place))
    return items


def standalone(it):
    cfg = (FftConfig.new(it["n"]).with_batch(it["batch"]).with_direction(it["direction"])
           .with_normalization(it["normalization"]))
    if it["algorithm"] == 2:
        cfg = cfg.with_tuning(FftTuning().with_force_bluestein_axes([0]))
    plan = FftPlan.c2c(cfg)
    d = to_dev(it["x"])
    out = empty_dev(it["x"].nbytes)
    plan.execute(d, out)
    return from_dev(out, np.complex64)


@pytest.mark.parametrize("inverse_mix", [False, True])
def test_batch_matches_standalone_bit_exact_and_fp64(torch_cuda, inverse_mix):
    import torch
    items = make_items(5 if inverse_mix else 4, 40, 6, inverse_mix)
    batch = FftBatch(items)
    info = batch.info()
    print("batch info", info, flush=True)
    batch.execute()
    torch.cuda.synchronize()
    worst = {}
    for it in items:
        got = from_dev(it["output"], np.complex64)
        ref = standalone(it)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), (it["n"], it["batch"])
        exp = c2c_expected(it["x"], [it["n"]], batch=it["batch"],
                           inverse=it["direction"] == FftDirection.Inverse, norm=int(it["normalization"]))
        e = oracle.rel_l2(got, exp)
        worst[it["n"]] = max(worst.get(it["n"], 0.0), e)
        assert e <= 1e-6, (it["n"], e)
    print("worst rel_l2 per length", worst, flush=True)


def test_batch_graph_replay(torch_cuda):
    """Captured once, replayed several times with eager work in between; every
    replay equals the eager execution."""
    import torch
    items = make_items(11, 24, 4)
    for it in items:  # out-of-place only, so the input survives replays
        if it["in_place"]:
            it["output"] = empty_dev(it["x"].nbytes)
    batch = FftBatch(items)
    batch.execute()
    torch.cuda.synchronize()
    eager = [from_dev(it["output"], np.complex64).copy() for it in items]
    s = torch.cuda.Stream()
    batch.capture(stream=s)
    for rep in range(3):
        for it in items:
            it["output"].zero_()
        torch.cuda.synchronize()
        batch.launch(stream=s)
        s.synchronize()
        for it, e in zip(items, eager):
            assert np.array_equal(from_dev(it["output"], np.complex64).view(np.uint32), e.view(np.uint32)), rep


@pytest.mark.parametrize("in_place", [False, True])
def test_batch_multipass_bluestein(torch_cuda, in_place):
    """Lengths whose Bluestein convolution exceeds the batched LDS budget run
    as chirp / padded FFT / filter / padded FFT / chirp phases (the padded
    65537-point transform is itself a four-step split); checked against the
    FP64 oracle, across batch counts, and for run-to-run determinism."""
    import torch
    cases = [(10007, 3), (5003, 1), (65537, 2), (2 * 3 * 1709, 1), (127, 4)]
    items = []
    for t, (n, b) in enumerate(cases):
        x = seeded_complex(30_000 + t, n * b)
        d_in = to_dev(x)
        d_out = d_in if in_place else empty_dev(x.nbytes)
        items.append(dict(n=n, batch=b, x=x, input=d_in, output=d_out,
                          direction=FftDirection.Forward, normalization=Normalization.Inverse))
    batch = FftBatch(items)
    info = batch.info()
    print("batch info", info, flush=True)
    assert info["phases"] >= 5
    batch.execute()
    torch.cuda.synchronize()
    first = [from_dev(it["output"], np.complex64).copy() for it in items]
    for it, got in zip(items, first):
        exp = c2c_expected(it["x"], [it["n"]], batch=it["batch"], inverse=False, norm=int(Normalization.Inverse))
        e = oracle.rel_l2(got, exp)
        print("n", it["n"], "batch", it["batch"], "rel_l2", e, flush=True)
        assert e <= 2e-6, (it["n"], e)
    if not in_place:
        batch.execute()
        torch.cuda.synchronize()
        for it, f in zip(items, first):
            assert np.array_equal(from_dev(it["output"], np.complex64).view(np.uint32), f.view(np.uint32))
