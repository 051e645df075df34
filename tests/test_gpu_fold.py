"""GPU tests of the 2D fold (plan.cpp fold_2d): axis 0 fused with the first
radix-2 stage of axis 1, then axis 1 as half-length lines over both halves.
Every compiled fold variant, forced through the tuning table, in both
directions, in place and out of place, against the FP64 oracle (relative
L2 <= 1e-6) and against the unfolded plan of the same shape."""
import numpy as np
import pytest

import oracle
from tests.helpers import c2c_expected, empty_dev, from_dev, seeded_complex, to_dev
import wgpufft_amd as W
from wgpufft_amd import FftConfig, FftPlan

pytestmark = pytest.mark.gpu

FOLD_VARIANTS = [
    "reg:f32:n4096:t256:l2:lf00:m0/0:p0:v2:16x16x16:fold",
    "reg:f32:n4096:t512:l2:lf00:m0/0:p0:v2:8x8x8x8:fold",
    "reg:f32:n4096:t512:l2:lf00:m0/0:p0:v2:8x8x8x8:pipe1:fold",
    "reg:f32:n4096:t512:l2:lf00:m0/0:p0:v2:8x8x8x8:pipe2:fold",
    "reg:f32:n4096:t256:l2:lf00:m0/0:p0:v2:16x16x16:pipe1:fold",
]


@pytest.fixture(autouse=True)
def _clean():
    W.clear_tuning()
    yield
    W.clear_tuning()


def fold_key(n0, lines, F=2):
    return f"f32|n{n0}|lf00|m0/0|p0|pair0|lines2^{(lines - 1).bit_length()}|fold{F}"


def _run(cfg, x, in_place):
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


@pytest.mark.parametrize("shape,batch", [([4096, 128], 3), ([4096, 2048], 1), ([4096, 8192], 1), ([2048, 2048], 1),
                                         ([2048, 512], 2), ([1024, 1024], 2), ([1024, 4096], 1)])
def test_fold_matches_oracle_and_unfolded(torch_cuda, shape, batch):
    n0, n1 = shape
    x = seeded_complex(500 + n1, n0 * n1 * batch)
    for inverse in (False, True):
        want = c2c_expected(x, shape, batch=batch, inverse=inverse)
        cfg = (FftConfig.inverse_nd(shape) if inverse else FftConfig.new_nd(shape)).with_batch(batch)
        W.clear_tuning()
        if n1 < 512:  # below M = 512 the fold runs only by table entry
            W.parse_tuning(f"fold|f32|n{n0}x{n1} = 2\n")
        plan, y = _run(cfg, x, in_place=inverse)
        desc = plan.describe()
        assert "fold2 M=%d" % n1 in desc, desc
        W.clear_tuning()
        e = oracle.rel_l2(y, want)
        W.parse_tuning(f"fold|f32|n{n0}x{n1} = 0\n")
        plain, y0 = _run(cfg, x, in_place=inverse)
        assert "fold" not in plain.describe(), plain.describe()
        e0 = oracle.rel_l2(y0, want)
        d = oracle.rel_l2(y, y0)
        print(f"{shape} b={batch} inverse={inverse} fold rel_l2={e:.3e} unfolded {e0:.3e} diff {d:.3e} [{desc}]",
              flush=True)
        assert e <= 1e-6 and e0 <= 1e-6 and d <= 1e-6


@pytest.mark.parametrize("variant", FOLD_VARIANTS)
def test_forced_fold_variants(torch_cuda, variant):
    shape, batch = [4096, 256], 5  # 640 row pairs: persistent forms loop over several tiles
    x = seeded_complex(77, 4096 * 256 * batch)
    W.parse_tuning(f"{fold_key(4096, 2 * 128 * batch)} = {variant}\nfold|f32|n4096x256 = 2\n")
    for inverse in (False, True):
        cfg = (FftConfig.inverse_nd(shape) if inverse else FftConfig.new_nd(shape)).with_batch(batch)
        plan, y = _run(cfg, x, in_place=False)
        desc = plan.describe()
        assert "fold" in desc and ("pipe" in desc) == (":pipe" in variant), (variant, desc)
        e = oracle.rel_l2(y, c2c_expected(x, shape, batch=batch, inverse=inverse))
        print(f"{variant} inverse={inverse} rel_l2={e:.3e}", flush=True)
        assert e <= 1e-6, (variant, inverse, e)


def test_fold_is_off_where_it_does_not_apply(torch_cuda):
    # Second axes of at least 512 points fold by default (the positive
    # control), so each guard below is the only thing keeping the fold off:
    # an odd second axis (also with a table entry asking for the fold), f64,
    # axes in the other order, a single axis of a 2D array.
    assert "fold2" in FftPlan.c2c(FftConfig.new_nd([4096, 512])).describe()
    assert "fold" not in FftPlan.c2c(FftConfig.new_nd([4096, 513])).describe()
    W.parse_tuning("fold|f32|n4096x3 = 2\nfold|f32|n4096x513 = 2\n")
    assert "fold" not in FftPlan.c2c(FftConfig.new_nd([4096, 3])).describe()
    assert "fold" not in FftPlan.c2c(FftConfig.new_nd([4096, 513])).describe()
    assert "fold" not in FftPlan.c2c(FftConfig.new_nd([4096, 512]).with_precision(W.FftPrecision.F64)).describe()
    assert "fold" not in FftPlan.c2c(FftConfig.new_nd([4096, 512]).with_precision(W.FftPrecision.Df64)).describe()
    assert "fold" not in FftPlan.c2c(FftConfig.new_nd([4096, 512]).with_axes([1, 0])).describe()
    assert "fold" not in FftPlan.c2c(FftConfig.new_nd([4096, 512]).with_axes([0])).describe()
    assert "fold" not in FftPlan.c2c(FftConfig.new_nd([4096, 512]).with_axes([1])).describe()


@pytest.mark.parametrize("shape,batch", [([4096, 256], 3), ([4096, 2048], 1)])
def test_fold4_matches_oracle(torch_cuda, shape, batch):
    """Radix-4 fold (table entry fold|f32|n<n0>x<M> = 4): rows r + b M/4 per
    thread, quarter-length columns."""
    n0, n1 = shape
    x = seeded_complex(900 + n1, n0 * n1 * batch)
    for inverse in (False, True):
        W.clear_tuning()
        W.parse_tuning(f"fold|f32|n{n0}x{n1} = 4\n")
        cfg = (FftConfig.inverse_nd(shape) if inverse else FftConfig.new_nd(shape)).with_batch(batch)
        plan, y = _run(cfg, x, in_place=not inverse)
        desc = plan.describe()
        assert "fold4 M=%d" % n1 in desc and "N=%d" % (n1 // 4) in desc, desc
        e = oracle.rel_l2(y, c2c_expected(x, shape, batch=batch, inverse=inverse))
        print(f"fold4 {shape} b={batch} inverse={inverse} rel_l2={e:.3e} [{desc}]", flush=True)
        assert e <= 1e-6, (shape, inverse, e)


@pytest.mark.parametrize("shape,batch", [([2048, 2048], 1), ([2048, 256], 3), ([2048, 128], 2), ([1024, 1024], 1),
                                         ([1024, 256], 2)])
def test_r2c_fold_matches_oracle(torch_cuda, shape, batch):
    """R2C fold (real rows r, r + M/2 in, folded packed rows out, then
    M/2-point packed columns) against the FP64 oracle, the unfolded plan and a
    C2R round trip."""
    import torch
    n0, n1 = shape
    xr = oracle.normal_f32(300 + n1 + n0, n0 * n1 * batch)
    want = oracle.r2c_nd(xr, shape, batch=batch, nthreads=8)
    outs = []
    for table in (f"fold|f32|r{n0}x{n1} = 2\n", f"fold|f32|r{n0}x{n1} = 0\n"):
        W.clear_tuning()
        W.parse_tuning(table)
        r2c = FftPlan.r2c(FftConfig.new_nd(shape).with_batch(batch))
        desc = r2c.describe()
        assert ("fold2" in desc) == table.endswith("2\n"), desc
        pk = empty_dev(r2c.required_output_buffer_size_bytes())
        r2c.execute(to_dev(xr), pk)
        torch.cuda.synchronize()
        outs.append(from_dev(pk, np.complex64))
        e = oracle.rel_l2(outs[-1], want)
        print(f"r2c {shape} b={batch} table={table.strip()!r} rel_l2={e:.3e} [{desc}]", flush=True)
        assert e <= 1e-6
    assert oracle.rel_l2(outs[0], outs[1]) <= 1e-6
    W.clear_tuning()
    c2r = FftPlan.c2r(FftConfig.inverse_nd(shape).with_batch(batch))
    back = empty_dev(xr.nbytes)
    c2r.execute(to_dev(outs[0]), back)
    torch.cuda.synchronize()
    assert oracle.rel_l2(from_dev(back, np.float32), xr) <= 1e-6


@pytest.mark.parametrize("shape,batch", [([2048, 2048], 1), ([2048, 256], 3), ([2048, 128], 2), ([1024, 1024], 1),
                                         ([1024, 256], 2)])
def test_c2r_unfold_matches_oracle(torch_cuda, shape, batch):
    """C2R unfold (half-length inverse columns over the even and odd rows,
    then the Hermitian-load axis-0 pass forming E_0 +- W_M^{-n} E_1) against
    the FP64 oracle and the plan without it, from an oracle spectrum whose
    self-conjugate bins carry imaginary parts (dropped by the C2R)."""
    import torch
    n0, n1 = shape
    P0 = n0 // 2 + 1
    pk = seeded_complex(400 + n1, P0 * n1 * batch)
    want = oracle.c2r_nd(pk, shape, batch=batch, nthreads=8)
    outs = []
    for table in (f"fold|f32|c{n0}x{n1} = 2\n", f"fold|f32|c{n0}x{n1} = 0\n"):
        W.clear_tuning()
        W.parse_tuning(table)
        c2r = FftPlan.c2r(FftConfig.inverse_nd(shape).with_batch(batch))
        desc = c2r.describe()
        assert ("unfold2" in desc) == table.endswith("2\n"), desc
        d = to_dev(pk)
        out = empty_dev(n0 * n1 * batch * 4)
        c2r.execute(d, out)
        torch.cuda.synchronize()
        outs.append(from_dev(out, np.float32))
        e = oracle.rel_l2(outs[-1], want)
        print(f"c2r {shape} b={batch} table={table.strip()!r} rel_l2={e:.3e} [{desc}]", flush=True)
        assert e <= 1e-6
        assert np.array_equal(from_dev(d, np.complex64), pk)  # the input survives
    assert oracle.rel_l2(outs[0], outs[1]) <= 1e-6


@pytest.mark.parametrize("norm", [W.Normalization.NONE, W.Normalization.Forward, W.Normalization.Orthogonal])
def test_fold_normalizations(torch_cuda, norm):
    """The scale lands on the fold's second pass (config.rs:262-286)."""
    shape = [1024, 512]
    x = seeded_complex(60 + int(norm), 1024 * 512)
    for inverse in (False, True):
        cfg = (FftConfig.inverse_nd(shape) if inverse else FftConfig.new_nd(shape)).with_normalization(norm)
        plan, y = _run(cfg, x, in_place=False)
        assert "fold2" in plan.describe(), plan.describe()
        e = oracle.rel_l2(y, c2c_expected(x, shape, inverse=inverse, norm=int(norm)))
        print(f"fold norm={norm.name} inverse={inverse} rel_l2={e:.3e}", flush=True)
        assert e <= 1e-6


def test_fold_recorder_graph_replay(torch_cuda):
    """Fold plans (C2C pair, R2C, C2R) recorded into one captured graph and
    replayed several times with eager work in between: every replay equals
    the separate executions bit for bit."""
    import torch
    from wgpufft_amd import FftRecorder
    shape = [2048, 512]
    x = seeded_complex(61, 2048 * 512)
    xr = oracle.normal_f32(62, 2048 * 512)
    fwd, inv = FftPlan.c2c(FftConfig.new_nd(shape)), FftPlan.c2c(FftConfig.inverse_nd(shape))
    r2c, c2r = FftPlan.r2c(FftConfig.new_nd(shape)), FftPlan.c2r(FftConfig.inverse_nd(shape))
    for p in (fwd, inv):
        assert "fold2" in p.describe(), p.describe()
    assert "fold2" in r2c.describe() and "unfold2" in c2r.describe(), (r2c.describe(), c2r.describe())
    a, b = to_dev(x), empty_dev(x.nbytes)
    ra, pk, back = to_dev(xr), empty_dev(r2c.required_output_buffer_size_bytes()), empty_dev(xr.nbytes)
    fwd.execute(a, b)
    inv.execute_in_place(b)
    r2c.execute(ra, pk)
    c2r.execute(pk, back)
    torch.cuda.synchronize()
    want_c, want_pk, want_r = (from_dev(b, np.complex64).copy(), from_dev(pk, np.complex64).copy(),
                               from_dev(back, np.float32).copy())
    s = torch.cuda.Stream()
    rec = FftRecorder(stream=s, capture_graph=True)
    fwd.record(rec, a, b)
    inv.record_in_place(rec, b)
    r2c.record(rec, ra, pk)
    c2r.record(rec, pk, back)
    rec.end()
    for i in range(3):
        b.zero_()
        pk.zero_()
        back.zero_()
        torch.cuda.synchronize()
        rec.launch(stream=s)
        s.synchronize()
        assert np.array_equal(from_dev(b, np.complex64).view(np.uint32), want_c.view(np.uint32)), i
        assert np.array_equal(from_dev(pk, np.complex64).view(np.uint32), want_pk.view(np.uint32)), i
        assert np.array_equal(from_dev(back, np.float32).view(np.uint32), want_r.view(np.uint32)), i


def test_fold_default_threshold(torch_cuda):
    """Without a table entry the fold runs from a second axis of 512 up."""
    assert "fold2" in FftPlan.c2c(FftConfig.new_nd([1024, 512])).describe()
    assert "fold2" not in FftPlan.c2c(FftConfig.new_nd([1024, 256])).describe()
    assert "fold2" in FftPlan.r2c(FftConfig.new_nd([2048, 512])).describe()
    assert "fold2" not in FftPlan.r2c(FftConfig.new_nd([2048, 256])).describe()
    assert "unfold2" in FftPlan.c2r(FftConfig.inverse_nd([2048, 512])).describe()
    assert "unfold2" not in FftPlan.c2r(FftConfig.inverse_nd([2048, 256])).describe()
