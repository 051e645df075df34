"""Host-memory plans (mifft_cpu_*, CpuFftPlan in src/cpu.rs:52-866) against
the FP64 oracle, over the reference's own CPU test configurations
(src/cpu.rs:914-1221) and bounds: f32 max error <= 2e-4 * peak and f64 / df64
<= 1e-12 * peak (cpu.rs:884-939, 942-985), plus relative L2 <= 1e-6 for f32;
the tiled, split-over-threads and four-step paths (cpu.rs:1165-1221) with
in place equal to out of place bit for bit; thread-count independence; and the
reference's length and error checks (cpu.rs:1067-1125, 1223-1237)."""
import numpy as np
import pytest

import oracle
import wgpufft_amd as W
from wgpufft_amd import CpuFftPlan, FftConfig, FftDirection, FftError, FftPrecision, Normalization


def signal(n, seed):
    """The reference tests' signal (cpu.rs:876-882)."""
    i = np.arange(n, dtype=np.float32)
    return ((np.sin(i * np.float32(0.37) + np.float32(seed)) + (np.arange(n) % 7).astype(np.float32) * 0.125)
            * 0.5).astype(np.float32)


def configs():
    c = [FftConfig.new(n).with_normalization(Normalization.NONE) for n in
         (1, 2, 3, 4, 5, 7, 8, 12, 16, 17, 29, 64, 100, 243)]
    c.append(FftConfig.inverse(96))
    c.append(FftConfig.new(60).with_normalization(Normalization.Orthogonal))
    c.append(FftConfig.new(30).with_normalization(Normalization.Forward))
    c.append(FftConfig.new_nd([4, 6]).with_batch(3))
    c.append(FftConfig.new_nd([8, 3, 5]).with_axes([2, 0]).with_batch(2))
    c.append(FftConfig.inverse_nd([6, 5, 4]).with_axes([1]))
    c.append(FftConfig.new_nd([5, 1, 7]).with_axes([0, 2]))
    return c


def expected(x, cfg):
    return oracle.c2c_nd(np.asarray(x, dtype=np.complex128), cfg.shape(), cfg.axes(), cfg.batch(),
                         cfg.direction() == FftDirection.Inverse, int(cfg.normalization()))


def total(cfg):
    return int(np.prod(cfg.shape())) * cfg.batch()


def close(actual, want, tol):
    peak = max(1.0, float(np.max(np.abs(want))))
    return float(np.max(np.abs(actual - want))) <= tol * peak


@pytest.mark.parametrize("i", range(len(configs())))
def test_c2c_f32_matches_oracle(i):
    cfg = configs()[i]
    x = signal(2 * total(cfg), 0.3)
    plan = CpuFftPlan.c2c(cfg)
    out = np.zeros(plan.required_output_len(), np.float32)
    plan.execute(x, out)
    want = expected(x.view(np.complex64), cfg)
    assert close(out.astype(np.float64), want.view(np.float64), 2e-4), cfg
    assert oracle.rel_l2(out.view(np.complex64), want) <= 1e-6


@pytest.mark.parametrize("i", range(len(configs())))
def test_c2c_f64_and_df64_match_oracle(i):
    cfg = configs()[i]
    v = signal(2 * total(cfg), 1.1).astype(np.float64)
    want = expected(v.view(np.complex128), cfg).view(np.float64)
    plan = CpuFftPlan.c2c(cfg.with_precision(FftPrecision.F64))
    out = np.zeros(plan.required_output_len(), np.float64)
    plan.execute_f64(v, out)
    assert close(out, want, 1e-12), cfg
    hilo = oracle.df64_split(v)
    plan = CpuFftPlan.c2c(cfg.with_precision(FftPrecision.Df64))
    o32 = np.zeros(plan.required_output_len(), np.float32)
    plan.execute(hilo, o32)
    assert close(oracle.df64_join(o32), want, 1e-12), cfg


def real_configs():
    return [FftConfig.new(1), FftConfig.new(2), FftConfig.new(9), FftConfig.new(16).with_batch(3),
            FftConfig.new_nd([6, 5]), FftConfig.new_nd([7, 4]).with_batch(2), FftConfig.new_nd([8, 3, 2]),
            FftConfig.new_nd([2048, 3]), FftConfig.new_nd([1025, 4]), FftConfig.new(1 << 16).with_batch(2)]


@pytest.mark.parametrize("i", range(len(real_configs())))
def test_r2c_matches_packed_oracle_f32_and_f64(i):
    cfg = real_configs()[i]
    real = signal(total(cfg), 0.7)
    want = oracle.r2c_nd(real.astype(np.float64), cfg.shape(), batch=cfg.batch())
    plan = CpuFftPlan.r2c(cfg)
    assert plan.packed_shape() == [cfg.shape()[0] // 2 + 1] + cfg.shape()[1:]
    packed = np.zeros(plan.required_output_len(), np.float32)
    plan.execute(real, packed)
    assert close(packed.astype(np.float64), want.view(np.float64), 2e-4)
    assert oracle.rel_l2(packed.view(np.complex64), want) <= 1e-6
    p64 = CpuFftPlan.r2c(cfg.with_precision(FftPrecision.F64))
    o64 = np.zeros(p64.required_output_len(), np.float64)
    p64.execute_f64(real.astype(np.float64), o64)
    assert close(o64, want.view(np.float64), 1e-12)


@pytest.mark.parametrize("i", range(len(real_configs())))
def test_c2r_matches_oracle_and_round_trips(i):
    cfg = real_configs()[i]
    real = signal(total(cfg), 2.3)
    fwd = CpuFftPlan.r2c(cfg)
    packed = np.zeros(fwd.required_output_len(), np.float32)
    fwd.execute(real, packed)
    packed[1] += 0.25  # the self-conjugate bins' imaginary parts are ignored
    inv_cfg = cfg.with_direction(FftDirection.Inverse)
    inv = CpuFftPlan.c2r(inv_cfg)
    back = np.zeros(inv.required_output_len(), np.float32)
    inv.execute(packed, back)
    want = oracle.c2r_nd(packed.view(np.complex64).astype(np.complex128), cfg.shape(), batch=cfg.batch())
    assert close(back.astype(np.float64), want, 2e-4)
    assert close(back.astype(np.float64), real.astype(np.float64), 2e-4)
    p64 = CpuFftPlan.c2r(inv_cfg.with_precision(FftPrecision.F64))
    b64 = np.zeros(p64.required_output_len(), np.float64)
    p64.execute_f64(packed.astype(np.float64), b64)
    assert close(b64, want, 1e-12)


def test_pure_tone_across_threads_is_one_delta():
    """cpu.rs:1026-1065: a complex exponential transforms to one scaled delta
    through the contiguous and tiled strided paths on several threads."""
    shape = [256, 128, 8]
    k = (5, 17, 3)
    i0, i1, i2 = np.meshgrid(np.arange(256), np.arange(128), np.arange(8), indexing="ij")
    ph = 2 * np.pi * (k[0] * i0 / 256 + k[1] * i1 / 128 + k[2] * i2 / 8)
    x = np.exp(1j * ph).transpose(2, 1, 0).reshape(-1)  # axis 0 contiguous
    plan = CpuFftPlan.c2c(FftConfig.new_nd(shape).with_precision(FftPrecision.F64)
                          .with_normalization(Normalization.Forward))
    out = np.zeros(2 * x.size, np.float64)
    plan.execute_f64(x.view(np.float64).copy(), out)
    y = out.view(np.complex128)
    peak = (k[2] * 128 + k[1]) * 256 + k[0]
    want = np.zeros_like(y)
    want[peak] = 1
    assert np.max(np.abs(y - want)) < 1e-9


def large_configs():
    return [FftConfig.new_nd([512, 400]), FftConfig.inverse_nd([300, 700]),
            FftConfig.new_nd([64, 48, 40]).with_batch(3), FftConfig.new_nd([40, 64, 48]).with_axes([2, 0]),
            FftConfig.new(100_000), FftConfig.inverse(98_304).with_batch(2),
            FftConfig.new_nd([70_000, 3]).with_axes([0]), FftConfig.new(65_537)]


@pytest.mark.parametrize("i", range(len(large_configs())))
def test_tiled_split_and_four_step_paths(i):
    """cpu.rs:1165-1221: f64 within 1e-12 of the peak and in place equal to
    out of place bit for bit; f32 within 2e-5 of the peak."""
    cfg = large_configs()[i]
    n = total(cfg)
    idx = np.arange(n, dtype=np.float64)
    vals = (np.sin(idx * 0.37) + 0.25) + 1j * (np.cos(idx * 0.11) - 0.5)
    want = expected(vals, cfg)
    peak = float(np.max(np.abs(want)))
    plan = CpuFftPlan.c2c(cfg.with_precision(FftPrecision.F64))
    inp = vals.view(np.float64).copy()
    out = np.zeros_like(inp)
    plan.execute_f64(inp, out)
    ip = inp.copy()
    plan.execute_in_place_f64(ip)
    assert np.array_equal(out, ip)
    assert np.max(np.abs(out.view(np.complex128) - want)) <= 1e-12 * peak
    p32 = CpuFftPlan.c2c(cfg)
    i32 = inp.astype(np.float32)
    o32 = np.zeros_like(i32)
    p32.execute(i32, o32)
    assert np.max(np.abs(o32.view(np.complex64).astype(np.complex128) - want)) <= 2e-5 * peak


@pytest.mark.parametrize("cfg", [FftConfig.new_nd([512, 400]), FftConfig.new(100_000),
                                 FftConfig.new_nd([64, 48, 40]).with_batch(3)])
def test_results_do_not_depend_on_thread_count(cfg):
    x = signal(2 * total(cfg), 0.9)
    outs = []
    for t in (1, 3, 0):
        W.set_cpu_threads(t)
        try:
            plan = CpuFftPlan.c2c(cfg)
            o = np.zeros_like(x)
            plan.execute(x, o)
            outs.append(o)
        finally:
            W.set_cpu_threads(0)
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


def test_lengths_and_errors():
    """cpu.rs:1067-1125, 1223-1237."""
    plan = CpuFftPlan.c2c(FftConfig.new(8).with_batch(2))
    assert plan.required_input_len() == 32
    with pytest.raises(FftError) as e:
        plan.execute(np.zeros(30, np.float32), np.zeros(32, np.float32))
    assert e.value.kind == "HostBufferLength"
    with pytest.raises(FftError) as e:
        plan.execute_f64(np.zeros(32), np.zeros(32))
    assert e.value.kind == "PrecisionUnsupported"
    assert CpuFftPlan.c2c(FftConfig.new(8).with_precision(FftPrecision.Df64)).required_input_len() == 32
    r2c = CpuFftPlan.r2c(FftConfig.new_nd([10, 3]))
    assert r2c.packed_shape() == [6, 3]
    assert (r2c.required_input_len(), r2c.required_output_len()) == (30, 36)
    for make, cfg, variant in [(CpuFftPlan.r2c, FftConfig.inverse(8), "InvalidRealTransformDirection"),
                               (CpuFftPlan.r2c, FftConfig.new_nd([4, 4]).with_axes([1]), "UnsupportedRealAxes"),
                               (CpuFftPlan.c2r, FftConfig.inverse(8).with_precision(FftPrecision.Df64),
                                "PrecisionUnsupported"),
                               (CpuFftPlan.c2c, FftConfig.new(0), "ZeroLength")]:
        with pytest.raises(FftError) as e:
            make(cfg)
        assert e.value.kind == variant, (cfg, e.value)
    with pytest.raises(FftError) as e:
        CpuFftPlan.r2c(FftConfig.new(8)).execute_in_place(np.zeros(8, np.float32))
    assert e.value.kind == "InPlaceUnsupported"
    p = CpuFftPlan.c2c(FftConfig.new(8))
    d = signal(16, 0.1)
    ref = np.zeros_like(d)
    p.execute(d, ref)
    p.execute_in_place(d)
    assert np.array_equal(d, ref)
