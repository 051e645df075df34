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


def test_every_table_entry_is_exercised():
    """Each entry of the committed table is taken by one of the plans above."""
    hit = set()
    plans = [FftPlan.c2c(FftConfig.new(1024).with_batch(4096)),
             FftPlan.c2c(FftConfig.new_nd([4096, 2048])), FftPlan.c2c(FftConfig.inverse_nd([4096, 2048])),
             FftPlan.r2c(FftConfig.new_nd([2048, 2048])), FftPlan.c2r(FftConfig.inverse_nd([2048, 2048])),
             FftPlan.c2c(FftConfig.new(1 << 22))]
    for shape in ([256, 256, 256], [1 << 24]):
        for p in (FftPrecision.F64, FftPrecision.Df64):
            plans.append(FftPlan.c2c(FftConfig.new_nd(shape).with_precision(p)))
            plans.append(FftPlan.c2c(FftConfig.inverse_nd(shape).with_precision(p)))
    entries = table_entries()
    for plan in plans:
        for li in range(plan.launch_count()):
            k, _ = plan.launch_tuning(li)
            if k in entries:
                hit.add(k)
    for n, prec in ((1 << 22, "f32"), (1 << 24, "f64"), (1 << 24, "df64")):
        hit.add(f"split|{prec}|n{n}")
    missing = sorted(set(entries) - hit)
    assert not missing, missing
