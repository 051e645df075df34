"""GPU parity of C2C plans against the FP64 oracle (reference_c2c_nd_f64).

Tolerances follow the reference's acceptance envelope: f32 relative L2 (its
rms_relative, tests/gpu_accuracy.rs:153-178) <= 1e-6 against FP64 and max-rel
well under 1e-5; f64 RMS <= 1e-13 (tests/gpu_f64.rs:15); df64 <= 1e-11
(tests/gpu_df64.rs:14).  Every call goes through the C ABI (libmifft.so).
"""
import numpy as np
import pytest

import oracle
from tests.helpers import c2c_expected, empty_dev, from_dev, seeded_complex, to_dev
from wgpufft_amd import (AxisKind, FftConfig, FftDirection, FftPlan, FftPrecision, FftRecorder,
                         FftTuning, Normalization)

pytestmark = pytest.mark.gpu

F32_L2 = 1e-6
F32_MAXREL = 5e-6


def run_c2c(cfg, x, in_place=False):
    plan = FftPlan.c2c(cfg)
    d_in = to_dev(x)
    if in_place:
        plan.execute_in_place(d_in)
        out = d_in
    else:
        out = empty_dev(plan.required_output_buffer_size_bytes())
        plan.execute(d_in, out)
    import torch
    torch.cuda.synchronize()
    dt = np.complex64 if cfg.precision() == FftPrecision.F32 else np.complex128
    return plan, from_dev(out, dt)


def check_f32(y, ref, l2=F32_L2, mr=F32_MAXREL, tag=""):
    e2, em = oracle.rel_l2(y, ref), oracle.max_rel(y, ref)
    print(f"{tag} rel_l2={e2:.3e} max_rel={em:.3e}", flush=True)
    assert e2 <= l2, (tag, e2)
    assert em <= mr, (tag, em)


LENGTHS_1D = [2, 3, 4, 5, 7, 8, 11, 13, 16, 17, 32, 64, 97, 120, 127, 128, 256, 343, 512, 1000,
              1001, 1009, 1024, 2048, 3256, 4096, 8192, 10007, 16384, 65536]


@pytest.mark.parametrize("n", LENGTHS_1D)
@pytest.mark.parametrize("direction", [FftDirection.Forward, FftDirection.Inverse])
def test_c2c_1d_lengths(torch_cuda, n, direction):
    batch = max(1, min(64, 200000 // n))
    x = seeded_complex(1000 + n, n * batch)
    cfg = FftConfig.new(n).with_batch(batch).with_direction(direction)
    plan, y = run_c2c(cfg, x)
    ref = c2c_expected(x, [n], batch=batch, inverse=direction == FftDirection.Inverse)
    check_f32(y, ref, tag=f"n={n} b={batch} {direction.name} [{plan.describe()}]")


def test_c2c_config1_1024x4096(torch_cuda):
    """1D N=1024, batch 4096, uniform[-1,1), seed 1 (the batched-line workload)."""
    n, b = 1024, 4096
    x = seeded_complex(1, n * b)
    plan, y = run_c2c(FftConfig.new(n).with_batch(b), x)
    assert plan.launch_count() == 1
    ref = c2c_expected(x, [n], batch=b)
    check_f32(y, ref, l2=5e-7, tag="C1")


def test_c2c_config2_4096x2048_forward_and_round_trip(torch_cuda):
    """2D [4096, 2048] in place forward then inverse; forward vs FP64, round trip vs input."""
    import torch
    shape = [4096, 2048]
    x = seeded_complex(2, 4096 * 2048)
    fwd = FftPlan.c2c(FftConfig.new_nd(shape))
    inv = FftPlan.c2c(FftConfig.inverse_nd(shape))
    assert fwd.launch_count() == 2 and fwd.supports_in_place()
    d = to_dev(x)
    fwd.execute_in_place(d)
    torch.cuda.synchronize()
    y = from_dev(d, np.complex64).copy()
    ref = c2c_expected(x, shape, nthreads=16)
    check_f32(y, ref, l2=5e-7, tag="C2 forward")
    inv.execute_in_place(d)
    torch.cuda.synchronize()
    z = from_dev(d, np.complex64)
    check_f32(z, x, l2=1e-6, mr=1e-5, tag="C2 round trip")


ND_CASES = [
    ([64, 48], None, 3),
    ([120, 7], None, 2),
    ([17, 32], None, 2),
    ([256, 256, 4], None, 1),
    ([8, 8, 8, 8], None, 2),
    ([2, 3], [1], 2),
    ([64, 30, 6], [0, 2], 2),
    ([30, 64, 6], [2, 1], 1),
    ([512, 256], [1], 1),
    ([1000, 10], [0], 4),
    ([16, 4096], [1], 1),
    ([1, 1], None, 1),
    ([127, 3, 5], None, 1),
    ([4, 65536], [1], 1),
]


@pytest.mark.parametrize("shape,axes,batch", ND_CASES)
@pytest.mark.parametrize("direction", [FftDirection.Forward, FftDirection.Inverse])
def test_c2c_nd(torch_cuda, shape, axes, batch, direction):
    cnt = int(np.prod(shape)) * batch
    x = seeded_complex(77 + cnt, cnt)
    cfg = FftConfig.new_nd(shape).with_batch(batch).with_direction(direction)
    if axes is not None:
        cfg = cfg.with_axes(axes)
    plan, y = run_c2c(cfg, x)
    ref = c2c_expected(x, shape, axes, batch, direction == FftDirection.Inverse)
    check_f32(y, ref, tag=f"{shape} axes={axes} b={batch} {direction.name} [{plan.describe()}]")


@pytest.mark.parametrize("norm", list(Normalization))
@pytest.mark.parametrize("direction", [FftDirection.Forward, FftDirection.Inverse])
def test_normalizations(torch_cuda, norm, direction):
    shape = [48, 20]
    x = seeded_complex(5, 960)
    cfg = FftConfig.new_nd(shape).with_normalization(norm).with_direction(direction)
    _, y = run_c2c(cfg, x)
    ref = c2c_expected(x, shape, inverse=direction == FftDirection.Inverse, norm=int(norm))
    check_f32(y, ref, tag=f"norm={norm.name} {direction.name}")


@pytest.mark.parametrize("shape", [[512, 256], [720, 480], [65536], [4241], [65537], [1009, 4], [96, 10007]])
def test_in_place_matches_out_of_place(torch_cuda, shape):
    """tests/gpu_in_place.rs: in-place equals out-of-place; here bit-exact (same kernels)."""
    cnt = int(np.prod(shape))
    x = seeded_complex(9, cnt)
    cfg = FftConfig.new_nd(shape)
    _, y_out = run_c2c(cfg, x)
    _, y_in = run_c2c(cfg, x, in_place=True)
    assert np.array_equal(y_out.view(np.uint32), y_in.view(np.uint32))


def test_recorder_bit_exact_and_graph_replay(torch_cuda):
    """tests/gpu_recorder.rs:87-124: recorded executions equal separate ones bit for bit."""
    import torch
    shape = [256, 128]
    x = seeded_complex(11, 256 * 128)
    fwd = FftPlan.c2c(FftConfig.new_nd(shape))
    inv = FftPlan.c2c(FftConfig.inverse_nd(shape))
    a = to_dev(x)
    b = empty_dev(fwd.required_output_buffer_size_bytes())
    fwd.execute(a, b)
    inv.execute_in_place(b)
    torch.cuda.synchronize()
    sep = from_dev(b, np.complex64).copy()
    a2 = to_dev(x)
    b2 = empty_dev(fwd.required_output_buffer_size_bytes())
    with FftRecorder() as rec:
        fwd.record(rec, a2, b2)
        inv.record_in_place(rec, b2)
    torch.cuda.synchronize()
    assert np.array_equal(from_dev(b2, np.complex64).view(np.uint32), sep.view(np.uint32))
    # graph capture on a side stream, replayed several times with eager work between
    s = torch.cuda.Stream()
    a3 = to_dev(x)
    b3 = empty_dev(fwd.required_output_buffer_size_bytes())
    torch.cuda.synchronize()
    rec = FftRecorder(stream=s, capture_graph=True)
    fwd.record(rec, a3, b3)
    inv.record_in_place(rec, b3)
    rec.end()
    for i in range(3):
        b3.zero_()
        torch.cuda.synchronize()
        rec.launch(stream=s)
        s.synchronize()
        got = from_dev(b3, np.complex64)
        assert np.array_equal(got.view(np.uint32), sep.view(np.uint32)), i


def test_forced_bluestein_and_rader(torch_cuda):
    """1009 forced Bluestein (FftTuning::with_force_bluestein_axes) and default Rader agree with FP64."""
    n, b = 1009, 8
    x = seeded_complex(3, n * b)
    ref = c2c_expected(x, [n], batch=b)
    p1, y1 = run_c2c(FftConfig.new(n).with_batch(b), x)
    assert p1.axis_kinds() == [AxisKind.Rader]
    check_f32(y1, ref, tag=f"rader [{p1.describe()}]")
    cfg = FftConfig.new(n).with_batch(b).with_tuning(FftTuning().with_force_bluestein_axes([0]))
    p2, y2 = run_c2c(cfg, x)
    assert p2.axis_kinds() == [AxisKind.Bluestein]
    check_f32(y2, ref, tag=f"bluestein [{p2.describe()}]")


def _f64_input(seed, cnt):
    r = oracle.uniform_f32(seed, 4 * cnt).astype(np.float64).reshape(-1, 2)
    # f64 values with full 53-bit mantissas: combine two f32 draws
    re = r[:, 0] + r[:, 1] * 2.0 ** -24
    r2 = oracle.uniform_f32(seed + 1, 2 * cnt).astype(np.float64)
    return (re[:cnt] + 1j * (r2[:cnt] + r2[cnt:2 * cnt] * 2.0 ** -24)).astype(np.complex128)


@pytest.mark.parametrize("shape,batch", [([256, 256, 8], 1), ([64, 64, 64], 1), ([1024], 16),
                                         ([127], 8), ([120, 9], 2), ([10007], 1)])
def test_f64(torch_cuda, shape, batch):
    cnt = int(np.prod(shape)) * batch
    x = _f64_input(21, cnt)
    cfg = FftConfig.new_nd(shape).with_batch(batch).with_precision(FftPrecision.F64)
    plan, y = run_c2c(cfg, x)
    ref = c2c_expected(x, shape, batch=batch)
    e = oracle.rel_l2(y, ref)
    print(f"f64 {shape} rel_l2={e:.3e} [{plan.describe()}]", flush=True)
    assert e <= 1e-13


@pytest.mark.parametrize("shape,batch", [([256, 256, 8], 1), ([64, 64, 64], 1), ([127], 8), ([120, 9], 2)])
def test_df64(torch_cuda, shape, batch):
    cnt = int(np.prod(shape)) * batch
    x = _f64_input(31, cnt)
    hilo = oracle.df64_split(x.view(np.float64))  # [re_hi, re_lo, im_hi, im_lo]
    cfg = FftConfig.new_nd(shape).with_batch(batch).with_precision(FftPrecision.Df64)
    plan = FftPlan.c2c(cfg)
    d = to_dev(hilo)
    out = empty_dev(plan.required_output_buffer_size_bytes())
    plan.execute(d, out)
    import torch
    torch.cuda.synchronize()
    y = oracle.df64_join(from_dev(out, np.float32)).view(np.complex128)
    xin = oracle.df64_join(hilo).view(np.complex128)
    ref = c2c_expected(xin, shape, batch=batch)
    e = oracle.rel_l2(y, ref)
    print(f"df64 {shape} rel_l2={e:.3e} [{plan.describe()}]", flush=True)
    assert e <= 1e-11


# Full-size high-precision configurations: 3D 256^3 (three axis passes) and the
# 1D 2^24 four-step, against the multithreaded FP64 oracle.
@pytest.mark.parametrize("shape,prec", [([256, 256, 256], "f64"), ([256, 256, 256], "df64"), ([1 << 24], "f64"),
                                        ([1 << 24], "df64")])
def test_high_precision_full_size(torch_cuda, shape, prec):
    import torch
    cnt = int(np.prod(shape))
    x = _f64_input(41, cnt)
    if prec == "f64":
        plan = FftPlan.c2c(FftConfig.new_nd(shape).with_precision(FftPrecision.F64))
        d = to_dev(x)
    else:
        plan = FftPlan.c2c(FftConfig.new_nd(shape).with_precision(FftPrecision.Df64))
        hilo = oracle.df64_split(x.view(np.float64))
        x = oracle.df64_join(hilo).view(np.complex128)
        d = to_dev(hilo)
    plan.execute_in_place(d)
    torch.cuda.synchronize()
    if prec == "f64":
        y = from_dev(d, np.complex128)
    else:
        y = oracle.df64_join(from_dev(d, np.float32)).view(np.complex128)
    ref = oracle.c2c_nd(x, shape, nthreads=16)
    e = oracle.rel_l2(y, ref)
    print(f"{prec} {shape} rel_l2={e:.3e} [{plan.describe()}]", flush=True)
    assert e <= (1e-13 if prec == "f64" else 1e-11)


# fuse_long_axes off (FftTuning::fuse_long_axes, tuning.rs:159-165): axes too
# long for one workgroup run as one Stockham pass per radix over HBM instead
# of the four-step split (axis_plan.rs:3540-3665).
@pytest.mark.parametrize("shape,batch,prec", [([65536], 3, "f32"), ([59049], 1, "f32"), ([4, 65536], 1, "f32"),
                                              ([1 << 18], 1, "f32"), ([65536], 2, "f64"), ([40000], 1, "df64")])
@pytest.mark.parametrize("direction", [FftDirection.Forward, FftDirection.Inverse])
def test_unfused_long_axes(torch_cuda, shape, batch, prec, direction):
    import torch
    cnt = int(np.prod(shape)) * batch
    tuning = FftTuning().with_fuse_long_axes(False)
    cfg = FftConfig.new_nd(shape).with_batch(batch).with_direction(direction).with_tuning(tuning)
    inverse = direction == FftDirection.Inverse
    if prec == "f32":
        x = seeded_complex(900 + cnt, cnt)
        plan, y = run_c2c(cfg, x, in_place=batch > 1)
        assert "stockham" in plan.describe(), plan.describe()
        check_f32(y, c2c_expected(x, shape, batch=batch, inverse=inverse), tag=f"unfused {shape} [{plan.describe()}]")
        return
    x = _f64_input(51, cnt)
    if prec == "f64":
        plan = FftPlan.c2c(cfg.with_precision(FftPrecision.F64))
        d = to_dev(x)
    else:
        plan = FftPlan.c2c(cfg.with_precision(FftPrecision.Df64))
        hilo = oracle.df64_split(x.view(np.float64))
        x = oracle.df64_join(hilo).view(np.complex128)
        d = to_dev(hilo)
    assert "stockham" in plan.describe(), plan.describe()
    plan.execute_in_place(d)
    torch.cuda.synchronize()
    y = from_dev(d, np.complex128) if prec == "f64" else oracle.df64_join(from_dev(d, np.float32)).view(np.complex128)
    e = oracle.rel_l2(y, c2c_expected(x, shape, batch=batch, inverse=inverse))
    print(f"unfused {prec} {shape} rel_l2={e:.3e} [{plan.describe()}]", flush=True)
    assert e <= (1e-13 if prec == "f64" else 1e-11)


# tests/gpu_accuracy.rs (:94-104): 4096, 3000, Rader 2999, Bluestein 3256 and a
# batched 2D, against FP64 with the reference's accuracy envelope.
@pytest.mark.parametrize("shape,batch,force", [([4096], 16, None), ([3000], 16, None), ([2999], 16, None),
                                               ([3256], 16, None), ([3256], 8, "bluestein"), ([120, 48], 6, None)])
def test_accuracy_envelope(torch_cuda, shape, batch, force):
    n = int(np.prod(shape))
    x = seeded_complex(1234 + n, n * batch)
    cfg = FftConfig.new_nd(shape).with_batch(batch)
    if force:
        cfg = cfg.with_tuning(FftTuning().with_force_bluestein_axes([0]))
    plan, y = run_c2c(cfg, x)
    ref = c2c_expected(x, shape, batch=batch)
    e2, em = oracle.rel_l2(y, ref), oracle.max_rel(y, ref)
    print(f"{shape} b={batch} {force} rms_rel={e2:.3e} max_rel={em:.3e} [{plan.describe()}]", flush=True)
    assert e2 < 5e-7, e2   # the reference's RMS-rel bound for non-fused routes
    assert em < 5e-7, em   # and its max-rel bound


def test_large_grid_dispatch(torch_cuda):
    """tests/gpu_dispatch_split.rs: a launch of more than 65535 workgroups (2^22
    + 64 lines of 64 points, 64 lines per workgroup).  Every line is one of 64
    seeded patterns; each must equal the FFT of its pattern bit for bit."""
    import torch
    n, lines = 64, (1 << 22) + 64
    pat = seeded_complex(77, n * 64).reshape(64, n)
    d = torch.from_numpy(pat.view(np.uint8).reshape(64, n * 8)).to("cuda").repeat(lines // 64, 1).reshape(-1)
    plan = FftPlan.c2c(FftConfig.new(n).with_batch(lines))
    plan.execute_in_place(d)
    torch.cuda.synchronize()
    y = d.view(torch.int32).reshape(lines // 64, 64 * n * 2)
    first = y[0:1]
    assert bool((y == first).all()), "a repeated line differs"
    got = from_dev(d[: 64 * n * 8], np.complex64).reshape(64, n)
    check_f32(got.reshape(-1), c2c_expected(pat.reshape(-1), [n], batch=64), tag=f"grid [{plan.describe()}]")


# tests/gpu_long_axes.rs: long 1D axes (two- and three-pass splits) and 1024^2,
# against FP64 with its envelope (max-rel < 2e-6, RMS-rel < 5e-7).
@pytest.mark.parametrize("shape,batch", [([8192], 4), ([1 << 17], 2), ([1 << 19], 1), ([1 << 21], 1),
                                         ([1024, 1024], 1)])
def test_long_axes_envelope(torch_cuda, shape, batch):
    n = int(np.prod(shape))
    x = seeded_complex(4321 + n, n * batch)
    plan, y = run_c2c(FftConfig.new_nd(shape).with_batch(batch), x)
    ref = c2c_expected(x, shape, batch=batch)
    e2, em = oracle.rel_l2(y, ref), oracle.max_rel(y, ref)
    print(f"{shape} b={batch} rms_rel={e2:.3e} max_rel={em:.3e} [{plan.describe()}]", flush=True)
    assert e2 < 5e-7 and em < 2e-6, (e2, em)


# tests/gpu_in_place.rs:22-245 configurations (length, batch, direction,
# forced Bluestein) beyond the shapes above
IN_PLACE_CASES = [([13], 1, False, False), ([17], 1, False, False), ([1523], 1, False, False),
                  ([47], 8, False, False), ([100, 47], 1, False, False), ([179, 179], 1, False, False),
                  ([256, 256, 3], 1, False, False), ([3, 8192], 1, False, False), ([31, 31, 31], 1, False, False),
                  ([47, 47], 1, False, False), ([60, 64], 1, True, False), ([64, 64], 1, False, False),
                  ([65536, 17], 1, False, False), ([65536, 3, 17], 1, False, False), ([11], 100, True, False),
                  ([2003], 64, True, False), ([97, 97, 97], 1, True, False), ([101], 1, False, True)]


@pytest.mark.parametrize("shape,batch,inverse,bluestein", IN_PLACE_CASES)
def test_in_place_reference_cases(torch_cuda, shape, batch, inverse, bluestein):
    cnt = int(np.prod(shape)) * batch
    x = seeded_complex(19, cnt)
    cfg = (FftConfig.inverse_nd(shape) if inverse else FftConfig.new_nd(shape)).with_batch(batch)
    if bluestein:
        cfg = cfg.with_tuning(FftTuning().with_force_bluestein_axes([0]))
    plan, y_out = run_c2c(cfg, x)
    _, y_in = run_c2c(cfg, x, in_place=True)
    ref = c2c_expected(x, shape, batch=batch, inverse=inverse)
    e_in = oracle.rel_l2(y_in, ref)
    print(f"in place {shape} b={batch} inverse={inverse} rel_l2={e_in:.3e} [{plan.describe()}]", flush=True)
    assert np.array_equal(y_out.view(np.uint32), y_in.view(np.uint32))
    assert e_in <= 1e-6
