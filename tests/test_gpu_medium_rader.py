"""GPU parity of the medium-prime Rader convolution (the reference's
medium_prime_rader_schedule and rader_axis_convolution,
src/runtime/rader_axis.rs:1185-1250; radices FUSED_PRIME_RADICES,
src/runtime/axis_plan.rs:2392): a prime N whose N - 1 has prime factors from
17 to 61 and at least 36 butterflies of the largest convolves cyclically over
N - 1 points with those radices, largest first.

The cases are the reference's own (tests/gpu_fused_prime.rs:141-181): 613,
strided 64 x 613, 1381, 4093, 4241 and 5003, with the schedules its unit test
pins (rader_axis.rs:2808-2820: 4241 -> 53 first, 1381 -> 23, 5003 -> 61, 41,
613 -> 17) and its four negatives (947, 283: too few butterflies; 1009: N - 1
smooth; 7727: a prime factor above 61).  Tolerances: the reference's per-bin
bound 5e-2 + 2e-5 |X| (gpu_fused_prime.rs:921-937) and relative L2 <= 1e-6
against the FP64 oracle."""
import numpy as np
import pytest

import oracle
from tests.helpers import c2c_expected, empty_dev, from_dev, seeded_complex, to_dev
from wgpufft_amd import FftBatch, FftConfig, FftDirection, FftPlan, Normalization

pytestmark = pytest.mark.gpu


def run(cfg, x):
    import torch
    plan = FftPlan.c2c(cfg)
    d = to_dev(x)
    out = empty_dev(x.nbytes)
    plan.execute(d, out)
    torch.cuda.synchronize()
    return plan, from_dev(out, np.complex64)


def check(y, ref, label):
    err = np.abs(y.astype(np.complex128) - ref)
    tol = 5e-2 + 2e-5 * np.abs(ref)
    assert np.all(np.isfinite(err)) and np.all(err <= tol), (label, float(np.max(err - tol)))
    e = oracle.rel_l2(y, ref)
    assert e <= 1e-6, (label, e)
    return e


# (label, shape, axes, batch, inverse, normalization, schedule prefix)
CASES = [
    ("613 batch 3", [613], None, 3, False, Normalization.Inverse, "17x"),
    ("613 inverse batch 3", [613], None, 3, True, Normalization.Inverse, "17x"),
    ("64x613 axis 1 batch 16", [64, 613], [1], 16, False, Normalization.NONE, "17x"),
    ("1381 inverse batch 2", [1381], None, 2, True, Normalization.Inverse, "23x"),
    ("1381 batch 1100", [1381], None, 1100, False, Normalization.Inverse, "23x"),  # two lines per tile
    ("1999 inverse batch 1030", [1999], None, 1030, True, Normalization.Inverse, "37x"),
    ("4093", [4093], None, 1, False, Normalization.Inverse, "31x"),
    ("4241 batch 2", [4241], None, 2, False, Normalization.Inverse, "53x16x5"),
    ("5003 inverse", [5003], None, 1, True, Normalization.Inverse, "61x41x2"),
    ("6841 batch 3", [6841], None, 3, False, Normalization.Inverse, "19x"),
    ("16x6841 axis 1", [16, 6841], [1], 1, True, Normalization.Inverse, "19x"),
]


@pytest.mark.parametrize("label,shape,axes,batch,inverse,norm,sched", CASES, ids=[c[0] for c in CASES])
def test_medium_prime_rader_matches_oracle(torch_cuda, label, shape, axes, batch, inverse, norm, sched):
    cfg = (FftConfig.inverse_nd(shape) if inverse else FftConfig.new_nd(shape)).with_batch(batch)
    cfg = cfg.with_normalization(norm)
    if axes is not None:
        cfg = cfg.with_axes(axes)
    p = shape[-1] if axes is not None else shape[0]
    x = seeded_complex(4100 + p + batch, int(np.prod(shape)) * batch)
    plan, y = run(cfg, x)
    assert f"rader p={p} M={p - 1} {sched}" in plan.describe(), plan.describe()
    if batch >= 1024 and axes is None and 1024 <= p - 1 <= 2400:
        assert "lpw=2" in plan.describe(), plan.describe()
    ref = c2c_expected(x, shape, axes=axes, batch=batch, inverse=inverse, norm=int(norm))
    e = check(y, ref, label)
    print(f"{label}: rel_l2={e:.3e} [{plan.describe()}]", flush=True)


@pytest.mark.parametrize("n", [947, 283, 1009, 7727])
def test_not_medium_prime(torch_cuda, n):
    # 946 = 22 * 43 and 282 = 6 * 47: too few butterflies; 1008 smooth
    # (plain cyclic Rader); 7726 = 2 * 3863 (zero-padded or Bluestein)
    plan = FftPlan.c2c(FftConfig.new(n).with_batch(2))
    d = plan.describe()
    assert f"M={n - 1} " not in d or n == 1009, d
    if n == 1009:
        assert "rader p=1009 M=1008 " in d, d
    x = seeded_complex(4200 + n, n * 2)
    _, y = run(FftConfig.new(n).with_batch(2), x)
    check(y, c2c_expected(x, [n], batch=2), str(n))


def test_medium_prime_batched_bit_exact(torch_cuda):
    # the persistent batched launch runs the same tile function (its medium
    # prime class): results bit-identical to standalone plans
    import torch
    items = []
    for t, (n, b) in enumerate([(613, 5), (1381, 2), (4241, 1), (613, 1), (120, 7)]):
        x = seeded_complex(4300 + t, n * b)
        d_in = to_dev(x)
        items.append(dict(n=n, batch=b, x=x, input=d_in, output=empty_dev(x.nbytes),
                          direction=FftDirection.Forward, normalization=Normalization.Inverse, algorithm=0,
                          in_place=False))
    batch = FftBatch(items)
    batch.execute()
    torch.cuda.synchronize()
    for it in items:
        got = from_dev(it["output"], np.complex64)
        _, want = run(FftConfig.new(it["n"]).with_batch(it["batch"]), it["x"])
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), it["n"]
        check(got, c2c_expected(it["x"], [it["n"]], batch=it["batch"]), str(it["n"]))
