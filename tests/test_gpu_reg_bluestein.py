"""GPU parity of the register-resident Bluestein kernel (the reference's
register Bluestein, src/runtime/register_fft.rs:510-664, bluestein_axis.rs:
937-969; Rader primes whose convolution would be zero-padded take it too,
rader_axis.rs:1262-1280): f32 Bluestein lengths and such Rader primes, every
power-of-two convolution length 64..8192, ragged line counts, both
directions, in place and out of place, against the oracle (relative L2 <=
1e-6, tests/gpu_accuracy.rs:94-104); strided axes of 2D and 3D shapes; forced
Rader keeps the Rader route."""
import numpy as np
import pytest

import oracle
from tests.helpers import c2c_expected, empty_dev, from_dev, seeded_complex, to_dev
from wgpufft_amd import FftBatch, FftConfig, FftDirection, FftPlan, FftTuning, Normalization

pytestmark = pytest.mark.gpu


def run(cfg, x, in_place=False):
    import torch
    plan = FftPlan.c2c(cfg)
    d = to_dev(x)
    if in_place:
        plan.execute_in_place(d)
        out = d
    else:
        out = empty_dev(x.nbytes)
        plan.execute(d, out)
    torch.cuda.synchronize()
    return plan, from_dev(out, np.complex64)


# (length, batch, forced Bluestein, convolution length)
CASES = [
    (31, 37, True, 64),      # forced: M = 64
    (34, 101, False, 128),   # 2 * 17: Bluestein, M = 128
    (37, 16, True, 128),
    (100, 5, True, 256),
    (131, 33, False, 512),   # prime, 130 = 2 * 5 * 13 smooth: forced below
    (263, 9, False, 1024),   # prime, 262 = 2 * 131: zero-padded Rader -> register Bluestein
    (1000, 3, True, 2048),
    (1229, 5, False, 4096),  # prime, 1228 = 4 * 307
    (3256, 3, False, 8192),  # 8 * 11 * 37: Bluestein (the reference's 1 GiB case)
    (4093, 2, True, 8192),   # prime, 4092 = 4 * 3 * 11 * 31: medium-prime Rader unless forced
]


@pytest.mark.parametrize("n,batch,forced,m", CASES)
@pytest.mark.parametrize("direction", [FftDirection.Forward, FftDirection.Inverse])
def test_register_bluestein_matches_oracle(torch_cuda, n, batch, forced, m, direction):
    if n == 131:
        forced = True  # cyclic Rader by default; the Bluestein route when forced
    tuning = FftTuning().with_force_bluestein_axes([0]) if forced else FftTuning()
    cfg = FftConfig.new(n).with_batch(batch).with_direction(direction).with_tuning(tuning)
    x = seeded_complex(900 + n + batch, n * batch)
    plan, y = run(cfg, x, in_place=batch % 2 == 1)
    assert f"bluestein-reg N={n} M={m} " in plan.describe(), plan.describe()
    ref = c2c_expected(x, [n], batch=batch, inverse=direction == FftDirection.Inverse)
    e = oracle.rel_l2(y, ref)
    print(f"N={n} b={batch} M={m} {direction.name} rel_l2={e:.3e} max_rel={oracle.max_rel(y, ref):.3e} "
          f"[{plan.describe()}]", flush=True)
    assert e <= 1e-6


def test_register_bluestein_normalization_and_2d(torch_cuda):
    # [1229, 6]: the contiguous prime axis through the register kernel, the
    # strided one as usual; orthogonal normalization (1/sqrt(N) each way)
    shape = [1229, 6]
    x = seeded_complex(77, 1229 * 6 * 2)
    cfg = FftConfig.new_nd(shape).with_batch(2).with_normalization(Normalization.Orthogonal)
    plan, y = run(cfg, x)
    assert "bluestein-reg N=1229 M=4096" in plan.describe(), plan.describe()
    ref = c2c_expected(x, shape, batch=2, norm=oracle.NORM_ORTHO)
    e = oracle.rel_l2(y, ref)
    print(f"{shape} orthogonal rel_l2={e:.3e} [{plan.describe()}]", flush=True)
    assert e <= 1e-6


def test_forced_rader_keeps_its_route(torch_cuda):
    p1 = FftPlan.c2c(FftConfig.new(1229).with_batch(2).with_tuning(FftTuning().with_force_rader_axes([0])))
    assert "bluestein-reg" not in p1.describe() and "rader" in p1.describe(), p1.describe()
    x = seeded_complex(78, 1229 * 2)
    _, y = run(FftConfig.new(1229).with_batch(2).with_tuning(FftTuning().with_force_rader_axes([0])), x)
    assert oracle.rel_l2(y, c2c_expected(x, [1229], batch=2)) <= 1e-6


# strided axes (the register kernel reads a line's points a row pitch apart):
# (shape, axes, batch, inverse, convolution length of the strided axis)
STRIDED = [
    ([8, 1229], [1], 3, False, 4096),
    ([1229, 1229], None, 1, True, 4096),
    ([5, 3256], [1], 2, False, 8192),
    ([33, 100], [1], 1, True, 256),      # Bluestein 100 = 4 * 25 forced below
    ([3, 263, 4], [1], 2, False, 1024),  # middle axis, stride 3
]


@pytest.mark.parametrize("shape,axes,batch,inverse,m", STRIDED)
def test_register_bluestein_strided_axes(torch_cuda, shape, axes, batch, inverse, m):
    cfg = (FftConfig.inverse_nd(shape) if inverse else FftConfig.new_nd(shape)).with_batch(batch)
    if axes is not None:
        cfg = cfg.with_axes(axes)
    if shape == [33, 100]:
        cfg = cfg.with_tuning(FftTuning().with_force_bluestein_axes([1]))
    n = shape[1]
    x = seeded_complex(990 + n + batch, int(np.prod(shape)) * batch)
    plan, y = run(cfg, x, in_place=batch % 2 == 1)
    assert f"bluestein-reg N={n} M={m} " in plan.describe(), plan.describe()
    ref = c2c_expected(x, shape, axes=axes, batch=batch, inverse=inverse)
    e = oracle.rel_l2(y, ref)
    print(f"{shape} axes={axes} b={batch} inverse={inverse} rel_l2={e:.3e} [{plan.describe()}]", flush=True)
    assert e <= 1e-6


def test_register_bluestein_batched_bit_exact(torch_cuda):
    # the persistent batched launch runs the 256-thread kernels; standalone
    # plans of M = 8192 run the 512-thread form of the same schedule:
    # results bit for bit the same
    import torch
    items = []
    for t, (n, b) in enumerate([(3256, 2), (4093, 1), (1229, 3), (34, 9)]):
        x = seeded_complex(950 + t, n * b)
        items.append(dict(n=n, batch=b, x=x, input=to_dev(x), output=empty_dev(x.nbytes),
                          direction=FftDirection.Forward, normalization=Normalization.Inverse,
                          algorithm=2 if n == 4093 else 0, in_place=False))
    batch = FftBatch(items)
    batch.execute()
    torch.cuda.synchronize()
    for it in items:
        got = from_dev(it["output"], np.complex64)
        cfg = FftConfig.new(it["n"]).with_batch(it["batch"])
        if it["algorithm"] == 2:
            cfg = cfg.with_tuning(FftTuning().with_force_bluestein_axes([0]))
        plan, want = run(cfg, it["x"])
        assert "bluestein-reg" in plan.describe(), plan.describe()
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), it["n"]
        assert oracle.rel_l2(got, c2c_expected(it["x"], [it["n"]], batch=it["batch"])) <= 1e-6
