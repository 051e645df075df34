"""The plans bench.py times, built with the committed MI355X tuning table
(wgpufft_amd/tuning/mi355x.txt) loaded as bench.py loads it, checked against
the FP64 oracle at full size:

* 1D C2C f32 N=1024 x 4096 forward (the pipelined 8x8x4x4 row kernel);
* 2D C2C f32 [4096, 2048] forward, inverse, and forward+inverse in place;
* 2D R2C / C2R f32 [2048, 2048] (fold and unfold passes);
* 3D C2C [256, 256, 256] and 1D [2^24] in f64 and df64;
* 1D C2C f32 [2^22] (the table's two-pass split entry).

Every launch whose tuning key the table holds must run the table's choice,
and every table entry must be taken by one of these plans.  Bounds: f32
relative L2 <= 1e-6 and max-rel <= 5e-6 against FP64 (tests/gpu_accuracy.rs:94-104),
f64 <= 1e-13 (tests/gpu_f64.rs:15), df64 <= 1e-11 (tests/gpu_df64.rs:14)."""
import os

import numpy as np
import pytest

import oracle
from tests.helpers import empty_dev, from_dev, seeded_complex, to_dev
import wgpufft_amd as W
from wgpufft_amd import FftConfig, FftPlan, FftPrecision

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TABLE = os.path.join(ROOT, "wgpufft_amd", "tuning", "mi355x.txt")


def table_entries():
    out = {}
    for line in open(TABLE):
        line = line.split("#")[0].strip()
        if "=" in line:
            k, v = line.split("=", 1)
            out[k.strip()] = v.strip()
    return out


@pytest.fixture(autouse=True)
def _table(torch_cuda):
    W.clear_tuning()
    W.load_tuning(TABLE)
    yield
    W.clear_tuning()


def _check_launches(plan):
    """Launches whose key the table holds run the table's variant."""
    entries = table_entries()
    for li in range(plan.launch_count()):
        k, v = plan.launch_tuning(li)
        if k in entries and entries[k].startswith("reg:"):
            assert v == entries[k], (k, v, entries[k])
    return plan


def test_c1_forward_1024x4096():
    import torch
    n, b = 1024, 4096
    x = seeded_complex(1, n * b)
    plan = _check_launches(FftPlan.c2c(FftConfig.new(n).with_batch(b)))
    assert "pipe2" in plan.describe(), plan.describe()
    d = to_dev(x)
    plan.execute_in_place(d)
    torch.cuda.synchronize()
    y = from_dev(d, np.complex64)
    ref = oracle.c2c_nd(x.astype(np.complex128), [n], batch=b, nthreads=16)
    e, m = oracle.rel_l2(y, ref), oracle.max_rel(y, ref)
    print(f"C1 table plan rel_l2={e:.3e} max_rel={m:.3e} [{plan.describe()}]", flush=True)
    assert e <= 1e-6 and m <= 5e-6


def test_c2_forward_inverse_4096x2048():
    import torch
    shape = [4096, 2048]
    x = seeded_complex(2, 4096 * 2048)
    fwd = _check_launches(FftPlan.c2c(FftConfig.new_nd(shape)))
    inv = _check_launches(FftPlan.c2c(FftConfig.inverse_nd(shape)))
    assert "fold2" in fwd.describe() and "fold2" in inv.describe()
    d = to_dev(x)
    fwd.execute_in_place(d)
    torch.cuda.synchronize()
    y = from_dev(d, np.complex64).copy()
    ref = oracle.c2c_nd(x.astype(np.complex128), shape, nthreads=16)
    e, m = oracle.rel_l2(y, ref), oracle.max_rel(y, ref)
    print(f"C2 forward rel_l2={e:.3e} max_rel={m:.3e} [{fwd.describe()}]", flush=True)
    assert e <= 1e-6 and m <= 5e-6
    # the inverse alone, against the oracle's inverse of the same input
    d2 = to_dev(x)
    inv.execute_in_place(d2)
    torch.cuda.synchronize()
    yi = from_dev(d2, np.complex64)
    refi = oracle.c2c_nd(x.astype(np.complex128), shape, inverse=True, nthreads=16)
    ei = oracle.rel_l2(yi, refi)
    print(f"C2 inverse rel_l2={ei:.3e} [{inv.describe()}]", flush=True)
    assert ei <= 1e-6
    # forward then inverse in place returns the input
    inv.execute_in_place(d)
    torch.cuda.synchronize()
    r = oracle.rel_l2(from_dev(d, np.complex64), x)
    print(f"C2 round trip rel_l2={r:.3e}", flush=True)
    assert r <= 1e-6


def test_r2c_c2r_2048x2048():
    import torch
    shape = [2048, 2048]
    xr = oracle.normal_f32(3, 2048 * 2048)
    r2c = _check_launches(FftPlan.r2c(FftConfig.new_nd(shape)))
    c2r = _check_launches(FftPlan.c2r(FftConfig.inverse_nd(shape)))
    d = to_dev(xr)
    pk = empty_dev(r2c.required_output_buffer_size_bytes())
    back = empty_dev(xr.nbytes)
    r2c.execute(d, pk)
    c2r.execute(pk, back)
    torch.cuda.synchronize()
    want = oracle.r2c_nd(xr, shape, nthreads=16)
    y = from_dev(pk, np.complex64)
    e1, m1 = oracle.rel_l2(y, want), oracle.max_rel(y, want)
    e2 = oracle.rel_l2(from_dev(back, np.float32), xr)
    # the C2R alone against the oracle's C2R of the same packed input
    pk2 = to_dev(y)
    back2 = empty_dev(xr.nbytes)
    c2r.execute(pk2, back2)
    torch.cuda.synchronize()
    e3 = oracle.rel_l2(from_dev(back2, np.float32), oracle.c2r_nd(y.astype(np.complex128), shape, nthreads=16))
    print(f"R2C rel_l2={e1:.3e} max_rel={m1:.3e} C2R(R2C) rel_l2={e2:.3e} C2R rel_l2={e3:.3e} "
          f"[{r2c.describe()} | {c2r.describe()}]", flush=True)
    assert e1 <= 1e-6 and m1 <= 5e-6 and e2 <= 1e-6 and e3 <= 1e-6


def _f64_input(seed, cnt):
    r = oracle.uniform_f32(seed, 2 * cnt).astype(np.float64)
    r2 = oracle.uniform_f32(seed + 1, 2 * cnt).astype(np.float64)
    return (r[0::2] + r2[0::2] * 2.0 ** -24) + 1j * (r[1::2] + r2[1::2] * 2.0 ** -24)


@pytest.mark.parametrize("shape", [[256, 256, 256], [1 << 24]])
@pytest.mark.parametrize("prec", ["f64", "df64"])
def test_high_precision_with_table(shape, prec):
    import torch
    cnt = int(np.prod(shape))
    x = _f64_input(5, cnt)
    p = FftPrecision.F64 if prec == "f64" else FftPrecision.Df64
    fwd = _check_launches(FftPlan.c2c(FftConfig.new_nd(shape).with_precision(p)))
    inv = _check_launches(FftPlan.c2c(FftConfig.inverse_nd(shape).with_precision(p)))
    if prec == "df64":
        hilo = oracle.df64_split(x.view(np.float64))
        x = oracle.df64_join(hilo).view(np.complex128)
        d = to_dev(hilo)
    else:
        d = to_dev(x)

    def read():
        if prec == "f64":
            return from_dev(d, np.complex128).copy()
        return oracle.df64_join(from_dev(d, np.float32)).view(np.complex128).copy()

    fwd.execute_in_place(d)
    torch.cuda.synchronize()
    y = read()
    ref = oracle.c2c_nd(x, shape, nthreads=16)
    e = oracle.rel_l2(y, ref)
    inv.execute_in_place(d)
    torch.cuda.synchronize()
    r = oracle.rel_l2(read(), x)
    print(f"{prec} {shape} forward rel_l2={e:.3e} round trip rel_l2={r:.3e} [{fwd.describe()}]", flush=True)
    tol = 1e-13 if prec == "f64" else 1e-11
    assert e <= tol and r <= tol


def test_f32_two_pass_entry_2p22():
    import torch
    n = 1 << 22
    x = seeded_complex(7, n)
    plan = _check_launches(FftPlan.c2c(FftConfig.new(n)))
    assert plan.describe().count("post") == 1, plan.describe()
    d = to_dev(x)
    out = empty_dev(x.nbytes)
    plan.execute(d, out)
    torch.cuda.synchronize()
    y = from_dev(out, np.complex64)
    e = oracle.rel_l2(y, oracle.c2c_nd(x.astype(np.complex128), [n], nthreads=16))
    print(f"2^22 f32 rel_l2={e:.3e} [{plan.describe()}]", flush=True)
    assert e <= 1e-6


def _entries_list():
    try:
        return sorted(table_entries().items())
    except OSError:
        return []


# configurations searched for keys the synthesis below does not reach: long
# 1D axes (three-pass splits) and the volumes of the published configurations
CANDIDATES = [[1 << k] for k in range(19, 25)] + [[128, 128, 64], [256, 256, 128], [1024, 1024], [2048, 2048],
                                                  [4096, 4096], [256, 256, 256]]


def _synthesize(key):
    """A plan configuration with a launch of tuning key `key`
    (prec|nN|lfAB|mL/S|pP|pairQ|lines2^K[|foldF]), its kind and extra table
    text (a forced split)."""
    f = key.split("|")
    prec, n = f[0], int(f[1][1:])
    lf, mm, post, pair = f[2][2:], f[3][1:], int(f[4][1:]), int(f[5][4:])
    lines = 1 << int(f[6][len("lines2^"):])
    fold = f[7] if len(f) > 7 else ""
    p = {"f32": FftPrecision.F32, "f64": FftPrecision.F64, "df64": FftPrecision.Df64}[prec]
    kind = {"0/0": "c2c", "1/1": "r2c", "2/2": "c2r"}[mm]
    if fold:
        return kind, [n, lines], None, 1, p, ""
    if lf == "00":
        return kind, [n], None, lines, p, ""
    if lf == "11" and not post:
        w = lines if pair else lines - 1  # an odd column count takes the one-line variants
        return kind, [w, n], [1], 1, p, ""
    # passes of a two-pass split of an n x n line: columns with the twiddle on
    # store (lf11, p1) and rows with the transposed store (lf01), n lines per line
    b = max(1, lines // n)
    return kind, [n * n], None, b, p, f"split|{prec}|n{n * n} = 2\n"


@pytest.mark.parametrize("key,choice", _entries_list())
def test_every_table_entry_runs_against_the_oracle(key, choice):
    """Each entry of the committed table: a plan with a launch of its key runs
    the table's choice and matches the oracle."""
    import torch
    if key.startswith("split|"):
        _, prec, nk = key.split("|")
        n = int(nk[1:])
        p = {"f32": FftPrecision.F32, "f64": FftPrecision.F64, "df64": FftPrecision.Df64}[prec]
        plan = FftPlan.c2c(FftConfig.new(n).with_precision(p))
        assert plan.describe().count("post") == int(choice) - 1, plan.describe()
        return
    if key.startswith("fold|"):
        return  # fold on/off entries: covered by the fold tests
    kind, shape, axes, batch, p, extra = _synthesize(key)
    if extra:
        W.parse_tuning(extra)

    def make(kind, shape, axes, batch, p):
        cfg = FftConfig.inverse_nd(shape) if kind == "c2r" else FftConfig.new_nd(shape)
        if axes:
            cfg = cfg.with_axes(axes)
        cfg = cfg.with_batch(batch).with_precision(p)
        plan = {"c2c": FftPlan.c2c, "r2c": FftPlan.r2c, "c2r": FftPlan.c2r}[kind](cfg)
        return plan, dict(plan.launch_tuning(i) for i in range(plan.launch_count()))
    plan, used = make(kind, shape, axes, batch, p)
    if key not in used:
        # passes of three-pass splits and of 3D volumes: the first candidate
        # configuration with a launch of this key (under the table alone)
        W.clear_tuning()
        W.load_tuning(TABLE)
        for cshape in CANDIDATES:
            cplan, cused = make(kind, cshape, None, 1, p)
            if key in cused:
                shape, axes, batch, plan, used = cshape, None, 1, cplan, cused
                break
    assert key in used, (key, list(used), plan.describe())
    assert used[key] == choice, (key, used[key], choice)
    cnt = int(np.prod(shape)) * batch
    if kind == "r2c":
        x = oracle.normal_f32(60, cnt)
        d = to_dev(x)
    elif kind == "c2r":
        x = oracle.r2c_nd(oracle.normal_f32(61, cnt), shape, batch=batch, nthreads=16).astype(np.complex64)
        d = to_dev(x)
    elif p == FftPrecision.F32:
        x = seeded_complex(62, cnt)
        d = to_dev(x)
    else:
        x = _f64_input(63, cnt)
        if p == FftPrecision.Df64:
            hilo = oracle.df64_split(x.view(np.float64))
            x = oracle.df64_join(hilo).view(np.complex128)
            d = to_dev(hilo)
        else:
            d = to_dev(x)
    out = empty_dev(plan.required_output_buffer_size_bytes())
    plan.execute(d, out)
    torch.cuda.synchronize()
    if kind == "r2c":
        y, want = from_dev(out, np.complex64), oracle.r2c_nd(x, shape, batch=batch, nthreads=16)
    elif kind == "c2r":
        y, want = from_dev(out, np.float32), oracle.c2r_nd(x.astype(np.complex128), shape, batch=batch, nthreads=16)
    elif p == FftPrecision.F32:
        y = from_dev(out, np.complex64)
        want = oracle.c2c_nd(x.astype(np.complex128), shape, axes, batch, nthreads=16)
    elif p == FftPrecision.F64:
        y, want = from_dev(out, np.complex128), oracle.c2c_nd(x, shape, axes, batch, nthreads=16)
    else:
        y = oracle.df64_join(from_dev(out, np.float32)).view(np.complex128)
        want = oracle.c2c_nd(x, shape, axes, batch, nthreads=16)
    e = oracle.rel_l2(y, want)
    tol = {FftPrecision.F32: 1e-6, FftPrecision.F64: 1e-13, FftPrecision.Df64: 1e-11}[p]
    print(f"{key} = {choice}: {kind} {shape} axes={axes} b={batch} rel_l2={e:.3e} [{plan.describe()}]", flush=True)
    assert e <= tol
