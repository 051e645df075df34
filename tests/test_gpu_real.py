"""GPU parity of R2C/C2R plans (packed [N0/2+1, N1, ...] spectrum, axis 0 halved)
against the FP64 oracle restating reference_r2c_packed_interleaved and
reference_c2r_from_packed_interleaved (src/math/mod.rs:81-153)."""
import numpy as np
import pytest

import oracle
from tests.helpers import empty_dev, from_dev, to_dev
from wgpufft_amd import FftConfig, FftError, FftPlan, FftPrecision

pytestmark = pytest.mark.gpu


def _sync():
    import torch
    torch.cuda.synchronize()


# the reference's own cases (tests/gpu_real.rs:22-120: 16, 17, 34, [17, 4],
# batch 5 of 16 and 17) and wider ones
REAL_SHAPES = [([4], 1), ([5], 3), ([16], 2), ([120], 4), ([1024], 8), ([2048, 2048], 1),
               ([64, 48], 2), ([30, 7], 1), ([256, 17, 3], 1), ([4096], 2), ([3, 2], 1),
               ([17], 1), ([34], 1), ([17, 4], 1), ([16], 5), ([17], 5)]


@pytest.mark.parametrize("shape,batch", REAL_SHAPES)
def test_r2c(torch_cuda, shape, batch):
    cnt = int(np.prod(shape)) * batch
    x = oracle.normal_f32(3, cnt)
    plan = FftPlan.r2c(FftConfig.new_nd(shape).with_batch(batch))
    assert plan.packed_shape() == [shape[0] // 2 + 1] + list(shape[1:])
    d = to_dev(x)
    out = empty_dev(plan.required_output_buffer_size_bytes())
    plan.execute(d, out)
    _sync()
    y = from_dev(out, np.complex64)
    ref = oracle.r2c_nd(x.astype(np.float64), shape, batch, nthreads=16)
    e2, em = oracle.rel_l2(y, ref), oracle.max_rel(y, ref)
    print(f"r2c {shape} b={batch} rel_l2={e2:.3e} max_rel={em:.3e} [{plan.describe()}]", flush=True)
    assert e2 <= 1e-6 and em <= 5e-6


@pytest.mark.parametrize("shape,batch", REAL_SHAPES)
def test_c2r_non_hermitian_input(torch_cuda, shape, batch):
    """Random packed input (not Hermitian-consistent): the C2R semantics of the
    reference (mirror, drop self-conjugate imaginary parts, real part) apply."""
    pk = int(np.prod(oracle.packed_shape(shape))) * batch
    p = oracle.uniform_f32(4, 2 * pk).view(np.complex64)
    plan = FftPlan.c2r(FftConfig.inverse_nd(shape).with_batch(batch))
    d = to_dev(p)
    out = empty_dev(plan.required_output_buffer_size_bytes())
    plan.execute(d, out)
    _sync()
    y = from_dev(out, np.float32)
    ref = oracle.c2r_nd(p.astype(np.complex128), shape, batch, nthreads=16)
    e2, em = oracle.rel_l2(y, ref), oracle.max_rel(y, ref)
    print(f"c2r {shape} b={batch} rel_l2={e2:.3e} max_rel={em:.3e} [{plan.describe()}]", flush=True)
    assert e2 <= 1e-6 and em <= 5e-6
    # the input buffer is left unchanged
    assert np.array_equal(from_dev(d, np.complex64).view(np.uint32), p.view(np.uint32))


def test_r2c_c2r_round_trip_2048(torch_cuda):
    shape = [2048, 2048]
    x = oracle.normal_f32(3, 2048 * 2048)
    r2c = FftPlan.r2c(FftConfig.new_nd(shape))
    c2r = FftPlan.c2r(FftConfig.inverse_nd(shape))
    d = to_dev(x)
    pk = empty_dev(r2c.required_output_buffer_size_bytes())
    back = empty_dev(c2r.required_output_buffer_size_bytes())
    r2c.execute(d, pk)
    c2r.execute(pk, back)
    _sync()
    z = from_dev(back, np.float32)
    e = oracle.rel_l2(z, x)
    print(f"r2c/c2r 2048^2 round trip rel_l2={e:.3e}", flush=True)
    assert e <= 1e-6


def test_real_in_place(torch_cuda):
    shape = [64, 10]
    x = oracle.normal_f32(8, 640)
    plan = FftPlan.r2c(FftConfig.new_nd(shape))
    assert not plan.supports_in_place()
    buf = empty_dev(plan.required_buffer_size_bytes())
    buf[: x.nbytes].copy_(to_dev(x))
    plan.execute_in_place(buf)
    _sync()
    y = from_dev(buf, np.complex64, 33 * 10)
    ref = oracle.r2c_nd(x.astype(np.float64), shape)
    assert oracle.rel_l2(y, ref) <= 1e-6


def test_real_precision_unsupported(torch_cuda):
    with pytest.raises(FftError) as e:
        FftPlan.r2c(FftConfig.new(16).with_precision(FftPrecision.F64))
    assert e.value.kind == "PrecisionUnsupported"


# Axis 0 that the fused real kernels cannot take: primes (Rader, Bluestein),
# composites with a large prime factor, and smooth lengths beyond one workgroup
# (four-step) run as widened complex lines (R2C) or Hermitian-extended lines
# (C2R) in the workspace.
GENERAL_SHAPES = [([17], 3), ([1009], 2), ([10007], 1), ([2 * 1709], 1), ([16384], 2), ([65536], 1),
                  ([17, 12], 1), ([1009, 6], 2), ([16384, 4], 1), ([23, 5, 3], 1), ([1 << 20], 1), ([34, 6], 2),
                  ([2 * 1009], 3)]


# Even axis 0 that is not fused runs at half length (N0/2-point complex lines
# and one twiddle pass); the table key fold|f32|h<N0> = 0 restores the
# widened N0-point lines.  Both against the oracle and each other.
@pytest.mark.parametrize("shape,batch", [([65536], 2), ([2 * 1709], 1), ([34, 6], 2), ([1 << 20], 1)])
def test_half_length_real_axis0(torch_cuda, shape, batch):
    import wgpufft_amd as W
    n0 = shape[0]
    cnt = int(np.prod(shape)) * batch
    x = oracle.normal_f32(15, cnt)
    pk = int(np.prod(oracle.packed_shape(shape))) * batch
    p = oracle.uniform_f32(16, 2 * pk).view(np.complex64)
    want_f = oracle.r2c_nd(x.astype(np.float64), shape, batch, nthreads=16)
    want_i = oracle.c2r_nd(p.astype(np.complex128), shape, batch, nthreads=16)
    got = {}
    try:
        for table in ("", f"fold|f32|h{n0} = 0\n"):
            W.clear_tuning()
            if table:
                W.parse_tuning(table)
            r2c = FftPlan.r2c(FftConfig.new_nd(shape).with_batch(batch))
            c2r = FftPlan.c2r(FftConfig.inverse_nd(shape).with_batch(batch))
            assert ("r2c post" in r2c.describe()) == (not table), r2c.describe()
            assert ("c2r pre" in c2r.describe()) == (not table), c2r.describe()
            out = empty_dev(r2c.required_output_buffer_size_bytes())
            r2c.execute(to_dev(x), out)
            back = empty_dev(c2r.required_output_buffer_size_bytes())
            c2r.execute(to_dev(p), back)
            _sync()
            yf, yi = from_dev(out, np.complex64), from_dev(back, np.float32)
            ef, ei = oracle.rel_l2(yf, want_f), oracle.rel_l2(yi, want_i)
            print(f"half={not table} {shape} b={batch} r2c {ef:.3e} c2r {ei:.3e} [{r2c.describe()}]", flush=True)
            assert ef <= 1e-6 and ei <= 1e-6
            got[bool(table)] = (yf, yi)
    finally:
        W.clear_tuning()
    assert oracle.rel_l2(got[False][0], got[True][0]) <= 1e-6
    assert oracle.rel_l2(got[False][1], got[True][1]) <= 1e-6


@pytest.mark.parametrize("shape,batch", GENERAL_SHAPES)
def test_r2c_general_axis0(torch_cuda, shape, batch):
    cnt = int(np.prod(shape)) * batch
    x = oracle.normal_f32(13, cnt)
    plan = FftPlan.r2c(FftConfig.new_nd(shape).with_batch(batch))
    d = to_dev(x)
    out = empty_dev(plan.required_output_buffer_size_bytes())
    plan.execute(d, out)
    _sync()
    y = from_dev(out, np.complex64)
    ref = oracle.r2c_nd(x.astype(np.float64), shape, batch, nthreads=16)
    e2, em = oracle.rel_l2(y, ref), oracle.max_rel(y, ref)
    print(f"r2c {shape} b={batch} rel_l2={e2:.3e} max_rel={em:.3e} [{plan.describe()}]", flush=True)
    assert e2 <= 1e-6 and em <= 5e-6


@pytest.mark.parametrize("shape,batch", GENERAL_SHAPES)
def test_c2r_general_axis0(torch_cuda, shape, batch):
    pk = int(np.prod(oracle.packed_shape(shape))) * batch
    p = oracle.uniform_f32(14, 2 * pk).view(np.complex64)
    plan = FftPlan.c2r(FftConfig.inverse_nd(shape).with_batch(batch))
    d = to_dev(p)
    out = empty_dev(plan.required_output_buffer_size_bytes())
    plan.execute(d, out)
    _sync()
    y = from_dev(out, np.float32)
    ref = oracle.c2r_nd(p.astype(np.complex128), shape, batch, nthreads=16)
    e2, em = oracle.rel_l2(y, ref), oracle.max_rel(y, ref)
    print(f"c2r {shape} b={batch} rel_l2={e2:.3e} max_rel={em:.3e} [{plan.describe()}]", flush=True)
    assert e2 <= 1e-6 and em <= 5e-6
    assert np.array_equal(from_dev(d, np.complex64).view(np.uint32), p.view(np.uint32))
