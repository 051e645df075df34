"""GPU parity of execution over logical views (FftPlan::execute_logical_views,
src/plan.rs:734-806; layouts src/runtime/logical_io.rs:33-178): strided,
offset and batch-strided inputs and outputs, against the FP64 oracle applied to
the logical data, with every element outside the output layout untouched."""
import numpy as np
import pytest

import oracle
from tests.helpers import c2c_expected, empty_dev, from_dev, seeded_complex, to_dev
from wgpufft_amd import FftConfig, FftLogicalLayout as L, FftPlan, FftPrecision, FftRecorder

pytestmark = pytest.mark.gpu


def positions(layout, per_batch, batch):
    bs = layout.batch_stride if layout.batch_stride else (per_batch - 1) * layout.element_stride + 1
    b = np.arange(batch)[:, None]
    i = np.arange(per_batch)[None, :]
    return (layout.element_offset + b * bs + i * layout.element_stride).reshape(-1)


def scatter_host(values, layout, per_batch, batch, fill, dtype):
    span = layout.required_element_span(per_batch, batch)
    buf = np.full(span + 3, fill, dtype=dtype)
    buf[positions(layout, per_batch, batch)] = values
    return buf


def run_views(plan, inbuf, lin, out_host, lout, workspace=False):
    import torch
    d_in = to_dev(inbuf)
    d_out = to_dev(out_host)
    if workspace:
        ws = empty_dev(max(16, plan.logical_workspace_size_bytes(lin, lout)))
        plan.execute_logical_views_with_workspace(d_in, lin, d_out, lout, ws)
    else:
        plan.execute_logical_views(d_in, lin, d_out, lout)
    torch.cuda.synchronize()
    return from_dev(d_out, out_host.dtype)


C2C_CASES = [
    # shape, batch, input layout, output layout
    ([1024], 5, L(3, 2, 2100), L(1, 3)),
    ([1000], 3, L(0, 1, 1007), L(2, 1)),          # batch-strided in, offset (unaligned) out
    ([64, 48], 2, L(1, 1), L(0, 1)),              # offset in (staged), contiguous out
    ([120, 9], 2, L(0, 4), L(5, 2, 2500)),        # Rader/mixed ND, both strided
    ([65536], 2, L(0, 2), L(0, 1, 65536 + 40)),   # four-step, strided in, batch-strided out
    ([1024, 512], 2, L(0, 2), L(0, 1, 1024 * 512 + 64)),  # 2D fold, strided in, batch-strided out
    ([2048, 512], 1, L(1, 1), L(3, 1)),           # 2D fold, offset in and out (staged)
]


@pytest.mark.parametrize("case", range(len(C2C_CASES)))
@pytest.mark.parametrize("workspace", [False, True])
def test_c2c_logical_views(torch_cuda, case, workspace):
    shape, batch, lin, lout = C2C_CASES[case]
    n = int(np.prod(shape))
    x = seeded_complex(500 + case, n * batch)
    plan = FftPlan.c2c(FftConfig.new_nd(shape).with_batch(batch))
    fill = np.complex64(7.0 - 3.0j)
    inbuf = scatter_host(x, lin, n, batch, np.complex64(1e6), np.complex64)
    outbuf = np.full(lout.required_element_span(n, batch) + 3, fill, dtype=np.complex64)
    got = run_views(plan, inbuf, lin, outbuf, lout, workspace)
    pos = positions(lout, n, batch)
    y = got[pos]
    ref = c2c_expected(x, shape, batch=batch)
    e = oracle.rel_l2(y, ref)
    print(f"{shape} b={batch} in={vars(lin)} out={vars(lout)} ws={workspace} rel_l2={e:.3e}", flush=True)
    assert e <= 1e-6
    mask = np.ones(got.size, bool)
    mask[pos] = False
    assert (got[mask] == fill).all(), "elements outside the output layout changed"


def test_f64_logical_views(torch_cuda):
    shape, batch = [256, 32], 3
    n = int(np.prod(shape))
    r = oracle.uniform_f32(77, 4 * n * batch).astype(np.float64).reshape(-1, 2)
    x = (r[:, 0] + 1j * r[:, 1])[: n * batch].astype(np.complex128)
    lin, lout = L(1, 3), L(0, 2, 2 * n + 11)
    plan = FftPlan.c2c(FftConfig.new_nd(shape).with_batch(batch).with_precision(FftPrecision.F64))
    inbuf = scatter_host(x, lin, n, batch, 0, np.complex128)
    outbuf = np.zeros(lout.required_element_span(n, batch) + 3, np.complex128)
    got = run_views(plan, inbuf, lin, outbuf, lout)
    e = oracle.rel_l2(got[positions(lout, n, batch)], c2c_expected(x, shape, batch=batch))
    print(f"f64 rel_l2={e:.3e}", flush=True)
    assert e <= 1e-13


def test_real_logical_views_round_trip(torch_cuda):
    """R2C from a strided real input into a batch-strided packed output, then
    C2R from that packed view into an offset real output."""
    import torch
    shape, batch = [32, 16], 3
    n = int(np.prod(shape))
    packed = (shape[0] // 2 + 1) * shape[1]
    real = oracle.uniform_f32(88, n * batch)
    lr, lp, lr2 = L(2, 3), L(1, 1, packed + 5), L(4, 1)
    r2c = FftPlan.r2c(FftConfig.new_nd(shape).with_batch(batch))
    c2r = FftPlan.c2r(FftConfig.inverse_nd(shape).with_batch(batch))
    d_real = to_dev(scatter_host(real, lr, n, batch, np.float32(9.0), np.float32))
    d_pack = to_dev(np.zeros(lp.required_element_span(packed, batch) + 3, np.complex64))
    r2c.execute_logical_views(d_real, lr, d_pack, lp)
    d_back = to_dev(np.zeros(lr2.required_element_span(n, batch) + 3, np.float32))
    c2r.execute_logical_views(d_pack, lp, d_back, lr2)
    torch.cuda.synchronize()
    spec = from_dev(d_pack, np.complex64)[positions(lp, packed, batch)]
    ref = oracle.r2c_nd(real.astype(np.float64), shape, batch)
    e = oracle.rel_l2(spec, ref)
    back = from_dev(d_back, np.float32)[positions(lr2, n, batch)]
    rt = float(np.linalg.norm(back - real) / np.linalg.norm(real))
    print(f"r2c rel_l2={e:.3e} round trip={rt:.3e}", flush=True)
    assert e <= 1e-6 and rt <= 1e-6


def test_in_place_strided_and_recorder(torch_cuda):
    """The same strided view as input and output transforms in place; a
    recorded logical execution gives the same bits as the direct one."""
    import torch
    shape, batch = [512], 4
    n = 512
    x = seeded_complex(99, n * batch)
    lay = L(1, 2)
    plan = FftPlan.c2c(FftConfig.new_nd(shape).with_batch(batch))
    host = scatter_host(x, lay, n, batch, np.complex64(5.0), np.complex64)
    d = to_dev(host)
    plan.execute_logical_views(d, lay, d, lay)
    d2 = to_dev(host)
    with FftRecorder() as rec:
        plan.record_logical_views(rec, d2, lay, d2, lay)
    torch.cuda.synchronize()
    a, b = from_dev(d, np.complex64), from_dev(d2, np.complex64)
    assert np.array_equal(a, b)
    pos = positions(lay, n, batch)
    assert oracle.rel_l2(a[pos], c2c_expected(x, shape, batch=batch)) <= 1e-6
    mask = np.ones(a.size, bool)
    mask[pos] = False
    assert (a[mask] == np.complex64(5.0)).all()


def test_out_of_bounds_rejected(torch_cuda):
    plan = FftPlan.c2c(FftConfig.new(64).with_batch(2))
    lay = L(0, 2)
    need = lay.required_element_span(64, 2) * 8
    d_in = empty_dev(need)
    d_out = empty_dev(need - 16)
    from wgpufft_amd import FftError
    with pytest.raises(FftError) as e:
        plan.execute_logical_views(d_in, lay, d_out, lay)
    assert e.value.kind == "BufferLayoutOutOfBounds"
