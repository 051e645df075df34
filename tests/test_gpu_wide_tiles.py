"""GPU parity of the 512-thread forms of the on-chip convolution and line
kernels: a fused Rader or Bluestein tile, or a runtime-schedule line, whose
LDS leaves room for one workgroup per CU runs 512 threads instead of 256,
the register Bluestein kernel of M = 8192 and the medium-prime Rader tiles
that share a CU run 512 threads at a 128-register budget.  Each case is
checked against the FP64 oracle (f32 relative L2 <= 1e-6,
tests/gpu_accuracy.rs:94-104; f64 <= 1e-13, tests/gpu_f64.rs:15; df64 <=
1e-11, tests/gpu_df64.rs:14), in both directions."""
import numpy as np
import pytest

import oracle
from tests.helpers import c2c_expected, empty_dev, from_dev, seeded_complex, to_dev
from wgpufft_amd import FftConfig, FftPlan, FftPrecision, FftTuning

pytestmark = pytest.mark.gpu

PREC = {"f32": FftPrecision.F32, "f64": FftPrecision.F64, "df64": FftPrecision.Df64}
TOL = {"f32": 1e-6, "f64": 1e-13, "df64": 1e-11}


def _f64_input(seed, cnt):
    r = oracle.uniform_f32(seed, 2 * cnt).astype(np.float64)
    r2 = oracle.uniform_f32(seed + 1, 2 * cnt).astype(np.float64)
    return (r[0::2] + r2[0::2] * 2.0 ** -24) + 1j * (r[1::2] + r2[1::2] * 2.0 ** -24)


def _run(cfg, prec, cnt, seed):
    import torch
    plan = FftPlan.c2c(cfg)
    if prec == "f32":
        x = seeded_complex(seed, cnt)
        d = to_dev(x)
    elif prec == "f64":
        x = _f64_input(seed, cnt)
        d = to_dev(x)
    else:
        x = _f64_input(seed, cnt)
        hilo = oracle.df64_split(x.view(np.float64))
        x = oracle.df64_join(hilo).view(np.complex128)
        d = to_dev(hilo)
    out = empty_dev(plan.required_output_buffer_size_bytes())
    plan.execute(d, out)
    torch.cuda.synchronize()
    if prec == "f32":
        y = from_dev(out, np.complex64)
    elif prec == "f64":
        y = from_dev(out, np.complex128)
    else:
        y = oracle.df64_join(from_dev(out, np.float32)).view(np.complex128)
    return plan, x, y


# (precision, length, batch, forced kind, route text)
CASES = [
    ("f32", 4099, 6, None, "rader p=4099 M=8232"),        # zero-padded Rader, 140 KB tile
    ("f32", 4106, 5, None, "bluestein N=4106 M=8232"),    # fused Bluestein, 140 KB tile
    ("f32", 6000, 7, None, "lds N=6000"),                 # runtime-schedule line, 102 KB tile
    ("f32", 3256, 3, None, "bluestein-reg N=3256 M=8192"),  # 512-thread register Bluestein
    ("f32", 4241, 3, None, "rader p=4241 M=4240 53x"),    # medium prime sharing a CU
    ("f32", 6841, 2, None, "rader p=6841 M=6840 19x"),    # medium prime, one tile per CU
    ("f64", 3001, 3, None, "rader p=3001 M=3000"),        # f64 Rader, 102 KB tile
    ("f64", 2053, 2, None, "rader p=2053 M=4116"),        # f64 zero-padded Rader, 140 KB tile
    ("f64", 3000, 3, None, "lds N=3000"),                 # f64 line, 102 KB tile
    ("f64", 1234, 3, None, "bluestein N=1234 M=2475"),    # f64 Bluestein, 84 KB tile
    ("df64", 3001, 2, None, "rader p=3001 M=3000"),
    ("df64", 1234, 2, None, "bluestein N=1234 M=2475"),
]


@pytest.mark.parametrize("prec,n,batch,forced,route", CASES)
@pytest.mark.parametrize("inverse", [False, True])
def test_wide_tiles_match_oracle(torch_cuda, prec, n, batch, forced, route, inverse):
    cfg = (FftConfig.inverse(n) if inverse else FftConfig.new(n)).with_batch(batch).with_precision(PREC[prec])
    plan, x, y = _run(cfg, prec, n * batch, 3000 + n + batch)
    assert route in plan.describe(), plan.describe()
    ref = c2c_expected(x, [n], batch=batch, inverse=inverse)
    e = oracle.rel_l2(y, ref)
    print(f"{prec} N={n} b={batch} inverse={inverse} rel_l2={e:.3e} [{plan.describe()}]", flush=True)
    assert e <= TOL[prec]


def test_wide_strided_rader_axis(torch_cuda):
    # a strided axis of a one-per-CU Rader tile (line-first loads, one line)
    shape = [6, 4099]
    cfg = FftConfig.new_nd(shape).with_axes([1]).with_tuning(FftTuning().with_force_rader_axes([1]))
    plan, x, y = _run(cfg, "f32", 6 * 4099, 3100)
    assert "rader p=4099 M=8232" in plan.describe(), plan.describe()
    e = oracle.rel_l2(y, c2c_expected(x, shape, axes=[1]))
    print(f"{shape} axis 1 rel_l2={e:.3e} [{plan.describe()}]", flush=True)
    assert e <= 1e-6
