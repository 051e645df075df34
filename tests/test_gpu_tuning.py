"""GPU tests of the autotuner: every registered alternative of a line shape
(radix split, threads per line, lines per workgroup) and the runtime-schedule
LDS shapes, forced through the tuning table, must transform correctly
(FP64 oracle, relative L2 <= 1e-6); mifft_autotune records a choice per pass
and plans built afterwards use it."""
import numpy as np
import pytest

import oracle
from tests.helpers import c2c_expected, empty_dev, from_dev, seeded_complex, to_dev
import wgpufft_amd as W
from wgpufft_amd import FftConfig, FftPlan

pytestmark = pytest.mark.gpu


def key(n, lf, pair, lines):
    lb = max(0, (lines - 1).bit_length())
    return f"f32|n{n}|lf{lf}|m0/0|p0|pair{int(pair)}|lines2^{lb}"


def reg(n, tpl, lines, lf, vec, rad):
    return f"reg:f32:n{n}:t{tpl}:l{lines}:lf{lf}:m0/0:p0:v{vec}:{rad}"


# (shape, batch, key of the pass forced, choices, text expected in describe())
CASES = [
    ([4096], 8, key(4096, "00", False, 8), [
        (reg(4096, 256, 1, "00", 1, "16x16x16"), "tpl=256"),
        (reg(4096, 512, 1, "00", 1, "8x8x8x8"), "8x8x8x8"),
        (reg(4096, 256, 2, "00", 1, "16x16x16"), "lpw=2"),
        (reg(4096, 128, 1, "00", 1, "16x16x16"), "tpl=128"),
        ("lds:lpw=1", "lds N=4096"), ("lds:lpw=2", "lpw=2")]),
    ([2048], 12, key(2048, "00", False, 12), [
        (reg(2048, 256, 1, "00", 1, "8x8x8x4"), "8x8x8x4"),
        (reg(2048, 128, 4, "00", 1, "16x16x8"), "lpw=4")]),
    ([1024], 20, key(1024, "00", False, 20), [
        (reg(1024, 128, 2, "00", 1, "8x8x4x4"), "8x8x4x4"),
        (reg(1024, 64, 8, "00", 1, "16x16x4"), "lpw=8"),
        ("lds:lpw=4", "lds N=1024")]),
    ([8, 4096], 1, key(4096, "11", True, 8), [
        (reg(4096, 256, 4, "11", 2, "16x16x16"), "vec=2 lpw=4")]),
    ([16, 2048], 1, key(2048, "11", True, 16), [
        (reg(2048, 128, 8, "11", 2, "16x16x8"), "vec=2 lpw=8"),
        (reg(2048, 256, 8, "11", 2, "8x8x8x4"), "8x8x8x4"),
        ("lds:lpw=16", "lds N=2048")]),
    ([32, 1024], 1, key(1024, "11", True, 32), [
        (reg(1024, 64, 16, "11", 2, "16x16x4"), "vec=2 lpw=16"),
        (reg(1024, 64, 8, "11", 2, "16x16x4"), "vec=2 lpw=8")]),
    # persistent pipelined forms, more tiles than workgroups (each loops)
    ([4096], 1200, key(4096, "00", False, 1200), [
        (reg(4096, 256, 1, "00", 1, "16x16x16") + ":pipe2", "tpl=256"),
        (reg(4096, 512, 1, "00", 1, "8x8x8x8") + ":pipe1", "8x8x8x8")]),
    ([4400, 2048], 1, key(2048, "11", True, 4400), [
        (reg(2048, 256, 4, "11", 2, "8x8x8x4") + ":pipe2", "vec=2 lpw=4")]),
    # strided line pairs in 256-thread workgroups (also the batched kernel's classes)
    ([32, 512], 1, key(512, "11", True, 32), [(reg(512, 32, 16, "11", 2, "8x8x8"), "vec=2 lpw=16")]),
    ([64, 256], 1, key(256, "11", True, 64), [(reg(256, 16, 32, "11", 2, "16x16"), "vec=2 lpw=32")]),
    ([128, 128], 1, key(128, "11", True, 128), [(reg(128, 8, 64, "11", 2, "16x8"), "vec=2 lpw=64")]),
    ([256, 64], 1, key(64, "11", True, 256), [(reg(64, 4, 128, "11", 2, "16x4"), "vec=2 lpw=128")]),
]


@pytest.fixture(autouse=True)
def _clean():
    W.clear_tuning()
    yield
    W.clear_tuning()


def _run(cfg, x):
    import torch
    plan = FftPlan.c2c(cfg)
    d = to_dev(x)
    out = empty_dev(x.nbytes)
    plan.execute(d, out)
    torch.cuda.synchronize()
    return plan, from_dev(out, np.complex64)


@pytest.mark.parametrize("case", range(len(CASES)))
def test_forced_alternatives_match_oracle(torch_cuda, case):
    shape, batch, k, choices = CASES[case]
    x = seeded_complex(70 + case, int(np.prod(shape)) * batch)
    ref = c2c_expected(x, shape, batch=batch)
    for choice, marker in choices:
        W.clear_tuning()
        W.parse_tuning(f"{k} = {choice}\n")
        for inverse in (False, True):
            cfg = (FftConfig.inverse_nd(shape) if inverse else FftConfig.new_nd(shape)).with_batch(batch)
            plan, y = _run(cfg, x)
            assert marker in plan.describe(), (choice, plan.describe())
            want = c2c_expected(x, shape, batch=batch, inverse=True) if inverse else ref
            e = oracle.rel_l2(y, want)
            print(f"{shape} b={batch} {choice} inverse={inverse} rel_l2={e:.3e}", flush=True)
            assert e <= 1e-6, (choice, inverse, e)


def test_autotune_records_and_plans_follow(torch_cuda, tmp_path):
    shape = [2048, 256]
    n, report = W.autotune(FftConfig.new_nd(shape), reps=5)
    print(report, flush=True)
    assert n >= 1 and W.tuning_size() >= 1
    assert report.count("->") == n
    x = seeded_complex(91, 2048 * 256)
    plan, y = _run(FftConfig.new_nd(shape), x)
    e = oracle.rel_l2(y, c2c_expected(x, shape))
    print(f"tuned {shape}: {plan.describe()} rel_l2={e:.3e}", flush=True)
    assert e <= 1e-6
    # the table survives a save / clear / load round trip and gives the same plan
    p = tmp_path / "t.txt"
    W.save_tuning(str(p))
    desc = plan.describe()
    W.clear_tuning()
    W.load_tuning(str(p))
    assert FftPlan.c2c(FftConfig.new_nd(shape)).describe() == desc


def test_autotune_sequence_pair_and_real_chain(torch_cuda):
    """A forward+inverse pair tuned as it runs (in place) and an R2C/C2R chain;
    plans built afterwards still match the oracle and round-trip."""
    from wgpufft_amd import FftTransformKind as K
    shape = [1024, 512]
    n, report = W.autotune_sequence([(FftConfig.new_nd(shape), K.C2c), (FftConfig.inverse_nd(shape), K.C2c)],
                                    reps=3)
    print(report, flush=True)
    assert n >= 1 and report.count("->") == n
    x = seeded_complex(92, 1024 * 512)
    plan, y = _run(FftConfig.new_nd(shape), x)
    assert oracle.rel_l2(y, c2c_expected(x, shape)) <= 1e-6
    n2, report2 = W.autotune_sequence([(FftConfig.new_nd([512, 256]), K.R2c),
                                       (FftConfig.inverse_nd([512, 256]), K.C2r)], reps=3)
    print(report2, flush=True)
    assert n2 >= 1
    import torch
    xr = oracle.normal_f32(93, 512 * 256)
    r2c = FftPlan.r2c(FftConfig.new_nd([512, 256]))
    c2r = FftPlan.c2r(FftConfig.inverse_nd([512, 256]))
    d = to_dev(xr)
    pk = empty_dev(r2c.required_output_buffer_size_bytes())
    back = empty_dev(xr.nbytes)
    r2c.execute(d, pk)
    c2r.execute(pk, back)
    torch.cuda.synchronize()
    assert oracle.rel_l2(from_dev(back, np.float32), xr) <= 1e-6


# Long 1D axes: the two-pass (n1 x n2) and three-pass (n3 x n1 x n2) splits,
# forced by the table's split entries, against the oracle in both directions.
@pytest.mark.parametrize("n,batch", [(1 << 18, 3), (1 << 19, 2), (1 << 20, 1)])
@pytest.mark.parametrize("split", ["2", "3"])
def test_forced_split_matches_oracle(torch_cuda, n, batch, split):
    x = seeded_complex(n + int(split), n * batch)
    W.parse_tuning(f"split|f32|n{n} = {split}\n")
    for inverse in (False, True):
        cfg = (FftConfig.inverse_nd([n]) if inverse else FftConfig.new_nd([n])).with_batch(batch)
        plan, y = _run(cfg, x)
        desc = plan.describe()
        assert desc.count("post") == int(split) - 1, desc
        want = c2c_expected(x, [n], batch=batch, inverse=inverse)
        e = oracle.rel_l2(y, want)
        print(f"n={n} b={batch} split={split} inverse={inverse} rel_l2={e:.3e} [{desc}]", flush=True)
        assert e <= 1e-6, (split, inverse, e)


# Default split of long f32 axes without table entries: three passes for one
# transform of 2^19 points and for 2^19 x 8 (2^22 points together), two
# passes from 2^23 points together (2^19 x 16), split factors of at most 2048.
@pytest.mark.parametrize("batch,passes", [(1, 3), (8, 3), (16, 2)])
def test_default_split_by_batch(torch_cuda, batch, passes):
    n = 1 << 19
    plan = FftPlan.c2c(FftConfig.new(n).with_batch(batch))
    desc = plan.describe()
    assert desc.count("post") == passes - 1, desc
    if batch == 16:
        x = seeded_complex(19 + batch, n * batch)
        _, y = _run(FftConfig.new(n).with_batch(batch), x)
        e = oracle.rel_l2(y, c2c_expected(x, [n], batch=batch))
        print(f"2^19 x {batch}: rel_l2={e:.3e} [{desc}]", flush=True)
        assert e <= 1e-6


# Persistent pipelined R2C / C2R axis-0 passes (real load + packed store,
# Hermitian-extended load + real store), more rows than workgroups so that
# each workgroup loops over several tiles, against the FP64 oracle.
@pytest.mark.parametrize("lines,minb", [(1, 2), (2, 2)])
def test_forced_pipelined_real_passes(torch_cuda, lines, minb):
    import torch
    n, batch = 2048, 1500
    lb = (batch - 1).bit_length()
    xr = oracle.normal_f32(94 + lines, n * batch)
    want = oracle.r2c_nd(xr, [n], batch=batch, nthreads=8)
    W.parse_tuning(
        f"f32|n{n}|lf00|m1/1|p0|pair0|lines2^{lb} = reg:f32:n{n}:t128:l{lines}:lf00:m1/1:p0:v1:16x16x8:pipe{minb}\n"
        f"f32|n{n}|lf00|m2/2|p0|pair0|lines2^{lb} = reg:f32:n{n}:t128:l{lines}:lf00:m2/2:p0:v1:16x16x8:pipe{minb}\n")
    r2c = FftPlan.r2c(FftConfig.new(n).with_batch(batch))
    c2r = FftPlan.c2r(FftConfig.inverse(n).with_batch(batch))
    assert "pipe" in r2c.describe() and "pipe" in c2r.describe(), (r2c.describe(), c2r.describe())
    d = to_dev(xr)
    pk = empty_dev(r2c.required_output_buffer_size_bytes())
    back = empty_dev(xr.nbytes)
    r2c.execute(d, pk)
    c2r.execute(pk, back)
    torch.cuda.synchronize()
    e1 = oracle.rel_l2(from_dev(pk, np.complex64), want)
    want_back = oracle.c2r_nd(want, [n], batch=batch, nthreads=8)
    e2 = oracle.rel_l2(from_dev(back, np.float32), want_back)
    print(f"pipelined real l{lines} pipe{minb}: r2c rel_l2={e1:.3e} c2r rel_l2={e2:.3e}", flush=True)
    assert e1 <= 1e-6 and e2 <= 1e-6
