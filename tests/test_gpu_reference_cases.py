"""The configurations of the reference's own GPU tests (tests/gpu_fused_pow2.rs,
gpu_fused_prime.rs, gpu_four_step.rs, gpu_small_volume.rs,
gpu_dispatch_split.rs, gpu_c2c.rs, gpu_long_axes.rs, gpu_f64.rs,
gpu_df64.rs), run through this engine's plans and checked against the FP64
oracle at the acceptance bounds (relative L2 <= 1e-6 for f32, 1e-13 for f64,
1e-11 for df64).  The reference compares its own kernels against each other and a
CPU plan; the values are what those comparisons pin."""
import numpy as np
import pytest

import oracle
from tests.helpers import c2c_expected, empty_dev, from_dev, seeded_complex, to_dev
from wgpufft_amd import FftConfig, FftError, FftPlan, FftPrecision, FftTuning, Normalization as Nm

pytestmark = pytest.mark.gpu

N, F, I, O = Nm.NONE, Nm.Forward, Nm.Inverse, Nm.Orthogonal

# (source, shape, batch, inverse, normalization, axes or None, tuning or None)
CASES = [
    # gpu_fused_pow2.rs
    ("fused_pow2", [2187], 1, False, N, None, None), ("fused_pow2", [2187], 1, True, I, None, None),
    ("fused_pow2", [256], 3, True, I, None, None), ("fused_pow2", [315], 3, True, I, None, None),
    ("fused_pow2", [4096], 1, False, N, None, None), ("fused_pow2", [1001], 1, False, N, None, None),
    ("fused_pow2", [13], 1, False, N, None, None), ("fused_pow2", [2048], 1, False, N, None, None),
    ("fused_pow2", [256, 12], 1, True, I, None, None), ("fused_pow2", [3, 256, 5], 1, False, I, None, None),
    ("fused_pow2", [3, 256, 5], 1, True, I, None, None), ("fused_pow2", [8, 16], 1, True, I, None, None),
    ("fused_pow2", [9, 25], 1, False, I, None, None),
    # gpu_fused_prime.rs
    ("fused_prime", [1024, 419], 1, True, I, [1], None), ("fused_prime", [1027, 1087], 1, True, I, [1], None),
    ("fused_prime", [2, 517, 3], 1, False, N, None, None), ("fused_prime", [6, 43, 5], 1, False, O, None, None),
    ("fused_prime", [6, 43, 5], 1, True, I, None, None), ("fused_prime", [1024, 986], 1, False, I, None, None),
    ("fused_prime", [2048, 179], 1, False, I, None, None), ("fused_prime", [64, 101], 1, False, I, None, None),
    ("fused_prime", [64, 613], 1, False, I, None, None), ("fused_prime", [769, 811], 1, False, I, None, None),
    ("fused_prime", [101], 1024, False, N, None, None), ("fused_prime", [101], 1024, True, I, None, None),
    ("fused_prime", [107], 1024, False, N, None, None), ("fused_prime", [1381], 2, True, I, None, None),
    ("fused_prime", [2062], 1, False, N, None, None), ("fused_prime", [2999], 1, True, I, None, None),
    ("fused_prime", [4093], 1, False, N, None, None), ("fused_prime", [4241], 2, False, N, None, None),
    ("fused_prime", [5003], 1, True, I, None, None), ("fused_prime", [613], 3, False, N, None, None),
    ("fused_prime", [613], 3, True, I, None, None),
    # gpu_four_step.rs
    ("four_step", [15, 14], 3, True, I, None, None), ("four_step", [15, 14], 1, False, N, None, None),
    ("four_step", [2, 101, 3], 1, True, I, None, None), ("four_step", [2, 101, 3], 1, False, N, None, None),
    ("four_step", [3, 5, 7, 11], 1, False, I, None, None), ("four_step", [3, 5, 7, 11], 1, True, I, None, None),
    ("four_step", [5, 17, 7], 3, True, I, None, None), ("four_step", [5, 7, 34], 2, True, I, None, None),
    ("four_step", [5, 7, 9], 3, True, I, None, None),
    # gpu_small_volume.rs
    ("small_volume", [17, 17], 5, True, I, None, None), ("small_volume", [64, 64, 32], 2, True, I, None, None),
    ("small_volume", [64, 64], 3, True, I, None, None), ("small_volume", [16, 16, 16], 1, False, N, None, None),
    ("small_volume", [16, 16, 256], 1, False, O, None, None), ("small_volume", [2, 101], 1, False, I, None, None),
    ("small_volume", [24, 20, 10], 1, False, N, None, None), ("small_volume", [31, 31], 1, False, N, None, None),
    ("small_volume", [31, 31], 1, False, I, None, FftTuning().with_direct_max_prime(0)),
    ("small_volume", [32, 16, 16], 1, False, N, None, None), ("small_volume", [47, 47], 1, False, I, None, None),
    ("small_volume", [51, 8, 3], 1, False, F, None, None), ("small_volume", [64, 64], 1, False, I, [1], None),
    ("small_volume", [8, 8, 8, 16], 1, False, N, None, None), ("small_volume", [8, 9, 10], 1, False, O, None, None),
    # gpu_c2c.rs
    ("c2c", [1], 1, True, I, None, None), ("c2c", [17], 1, True, I, None, None), ("c2c", [34], 1, True, I, None, None),
    ("c2c", [64], 1, True, I, None, None), ("c2c", [8], 1, True, I, None, None), ("c2c", [17, 4], 2, True, I, None, None),
    ("c2c", [2, 3], 2, True, I, None, None), ("c2c", [34, 16], 1, True, I, None, None),
    ("c2c", [4, 3], 1, True, I, [0], None), ("c2c", [1], 1, False, N, None, None),
    ("c2c", [1024], 1, False, N, None, None), ("c2c", [16], 1, False, N, None, None),
    ("c2c", [17], 2, False, I, None, None), ("c2c", [34], 1, False, N, None, None),
    ("c2c", [64], 1, False, F, None, None), ("c2c", [8], 1, False, N, None, None),
    ("c2c", [16, 16], 1, False, N, None, None), ("c2c", [17, 3, 2], 1, False, N, None, None),
    ("c2c", [17, 34], 1, False, N, None, None), ("c2c", [2, 64], 1, False, I, None, None),
    ("c2c", [34, 4], 1, False, I, None, None), ("c2c", [4, 17], 1, False, N, None, None),
    ("c2c", [4, 3, 2], 1, False, N, None, None), ("c2c", [4, 34], 1, False, N, None, None),
    ("c2c", [64, 2], 1, False, N, None, None), ("c2c", [64, 64], 1, False, N, None, None),
    # gpu_long_axes.rs
    ("long_axes", [12288], 1, True, I, None, None), ("long_axes", [32768], 1, True, I, None, None),
    ("long_axes", [3, 65536], 1, True, I, [1], None), ("long_axes", [5, 4096], 1, True, I, [1], None),
    ("long_axes", [8192, 4], 1, True, I, None, None), ("long_axes", [1 << 21], 1, False, N, None, None),
    ("long_axes", [16384], 1, False, I, None, None), ("long_axes", [65536], 1, False, N, None, None),
    ("long_axes", [8192], 1, False, N, None, None), ("long_axes", [1024, 1024], 1, False, I, None, None),
    ("long_axes", [16, 2048], 1, False, I, None, None), ("long_axes", [3, 10000, 2], 1, False, I, None, None),
    ("long_axes", [3, 8192], 1, False, I, None, None), ("long_axes", [65536, 3], 1, False, I, None, None),
    ("long_axes", [8192, 2], 1, False, I, None, None),
    # gpu_dispatch_split.rs: a Bluestein line count above 65535 workgroups
    ("dispatch_split", [34], 62601, False, I, None, None), ("dispatch_split", [34], 62601, True, I, None, None),
]


@pytest.mark.parametrize("case", range(len(CASES)), ids=[f"{c[0]}-{c[1]}x{c[2]}{'-inv' if c[3] else ''}"
                                                          for c in CASES])
def test_reference_configuration(torch_cuda, case):
    import torch
    src, shape, batch, inverse, norm, axes, tuning = CASES[case]
    cfg = (FftConfig.inverse_nd(shape) if inverse else FftConfig.new_nd(shape)).with_batch(batch)
    cfg = cfg.with_normalization(norm)
    if axes is not None:
        cfg = cfg.with_axes(axes)
    if tuning is not None:
        cfg = cfg.with_tuning(tuning)
    cnt = int(np.prod(shape)) * batch
    x = seeded_complex(1000 + case, cnt)
    plan = FftPlan.c2c(cfg)
    d = to_dev(x)
    out = empty_dev(plan.required_output_buffer_size_bytes())
    plan.execute(d, out)
    torch.cuda.synchronize()
    y = from_dev(out, np.complex64)
    ref = c2c_expected(x, shape, axes=axes, batch=batch, inverse=inverse, norm=int(norm))
    e = oracle.rel_l2(y, ref)
    print(f"{src} {shape} b={batch} inverse={inverse} norm={norm.name} axes={axes} rel_l2={e:.3e} "
          f"[{plan.describe()}]", flush=True)
    assert e <= 1e-6


# gpu_f64.rs and gpu_df64.rs: (shape, batch, inverse, normalization)
HIGH = [([1], 1, True, I), ([16], 1, True, I), ([4096], 1, True, I), ([45], 1, True, I), ([60], 1, True, I),
        ([8, 15], 1, True, I), ([1], 1, False, I), ([16], 1, False, I), ([2048], 1, False, I),
        ([3000], 1, False, I), ([45], 1, False, I), ([60], 1, False, I), ([64], 1, False, N), ([8], 1, False, I),
        ([8, 15], 1, False, I), ([2], 1, True, I), ([34], 1, True, I), ([67], 1, True, I), ([3, 17], 1, True, I),
        ([9, 10], 1, True, I), ([17], 1, False, I), ([68], 1, False, I)]


def _f64_values(seed, cnt):
    a = oracle.uniform_f32(seed, 4 * cnt).astype(np.float64).reshape(4, cnt)
    return (a[0] + a[1] * 2.0 ** -24) + 1j * (a[2] + a[3] * 2.0 ** -24)


@pytest.mark.parametrize("prec", ["f64", "df64"])
@pytest.mark.parametrize("case", range(len(HIGH)), ids=[f"{c[0]}{'-inv' if c[2] else ''}" for c in HIGH])
def test_reference_high_precision(torch_cuda, prec, case):
    import torch
    shape, batch, inverse, norm = HIGH[case]
    cfg = (FftConfig.inverse_nd(shape) if inverse else FftConfig.new_nd(shape)).with_batch(batch)
    cfg = cfg.with_normalization(norm).with_precision(FftPrecision.F64 if prec == "f64" else FftPrecision.Df64)
    cnt = int(np.prod(shape)) * batch
    x = _f64_values(2000 + case, cnt).astype(np.complex128)
    plan = FftPlan.c2c(cfg)
    if prec == "f64":
        d = to_dev(x)
    else:
        hilo = oracle.df64_split(x.view(np.float64))
        x = oracle.df64_join(hilo).view(np.complex128)
        d = to_dev(hilo)
    out = empty_dev(plan.required_output_buffer_size_bytes())
    plan.execute(d, out)
    torch.cuda.synchronize()
    y = from_dev(out, np.complex128) if prec == "f64" else oracle.df64_join(from_dev(out, np.float32)).view(np.complex128)
    ref = c2c_expected(x, shape, batch=batch, inverse=inverse, norm=int(norm))
    e = oracle.rel_l2(y, ref)
    print(f"{prec} {shape} inverse={inverse} rel_l2={e:.3e} [{plan.describe()}]", flush=True)
    assert e <= (1e-13 if prec == "f64" else 1e-11)


def test_reference_invalid_axis(torch_cuda):
    """gpu_c2c.rs: FftConfig::new(8).with_axes([1]) is InvalidAxis."""
    with pytest.raises(FftError) as e:
        FftPlan.c2c(FftConfig.new(8).with_axes([1]))
    assert e.value.kind == "InvalidAxis"


@pytest.mark.parametrize("shape,batch", [([30, 12], 1), ([64, 32], 1), ([15, 14], 1), ([8, 8, 32], 3)])
def test_reference_recorder_shapes(torch_cuda, shape, batch):
    """gpu_recorder.rs shapes: forward (no normalization) and inverse recorded
    on one stream equal separate executions bit for bit and the oracle."""
    import torch
    from wgpufft_amd import FftRecorder
    cnt = int(np.prod(shape)) * batch
    x = seeded_complex(3000 + cnt, cnt)
    fwd = FftPlan.c2c(FftConfig.new_nd(shape).with_batch(batch).with_normalization(N))
    inv = FftPlan.c2c(FftConfig.inverse_nd(shape).with_batch(batch))
    a, b = to_dev(x), empty_dev(x.nbytes)
    fwd.execute(a, b)
    torch.cuda.synchronize()
    mid = from_dev(b, np.complex64).copy()
    inv.execute_in_place(b)
    torch.cuda.synchronize()
    sep = from_dev(b, np.complex64).copy()
    assert oracle.rel_l2(mid, c2c_expected(x, shape, batch=batch, norm=int(N))) <= 1e-6
    assert oracle.rel_l2(sep, c2c_expected(mid, shape, batch=batch, inverse=True)) <= 1e-6
    a2, b2 = to_dev(x), empty_dev(x.nbytes)
    with FftRecorder() as rec:
        fwd.record(rec, a2, b2)
        inv.record_in_place(rec, b2)
    torch.cuda.synchronize()
    assert np.array_equal(from_dev(b2, np.complex64).view(np.uint32), sep.view(np.uint32))
