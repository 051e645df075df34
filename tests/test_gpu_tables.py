"""The device tables as tables: twiddles, stage twiddles, chirps, Rader and
Bluestein filter spectra and permutations, four-step twiddle factors, copied
back from the GPU (mifft_twiddle_table, mifft_plan_table) and checked entry by
entry against mpmath at 40 digits.

* f32 twiddles and chirps are the exact values rounded once: bit-equal to the
  committed goldens (tests/golden/twiddle_f32.json, chirp_f32.json, made by
  make_golden.py with once-rounding; the reference asserts this for its host
  LUT, src/runtime/twiddle.rs:198-237) and to mpmath over whole tables;
* f64 tables are within 1 ulp of the exact values (twiddle_f64.json);
* filter spectra FFT_m(b)/m (bluestein_axis.rs:818-856, rader_axis.rs:1318-1336)
  are computed in FP64 on the device and rounded once: against numpy's FP64
  FFT of the mpmath sequence, within one f32 ulp (f64: 1e-14 relative to the
  largest entry);
* Rader permutations are g^q mod p and g^-q mod p exactly (rader_axis.rs:1291-1304).
"""
import json
import os
import struct

import numpy as np
import pytest

import wgpufft_amd as W
from wgpufft_amd import FftConfig, FftPlan, FftPrecision, FftTuning

mp = pytest.importorskip("mpmath")
mp.mp.dps = 40

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def exact(t, n):
    """exp(-2 pi i t / n) at 40 digits."""
    return mp.expjpi(-mp.mpf(2 * (t % n)) / n)


def once_f32(v):
    with mp.workprec(24):
        return np.float32(float(+mp.mpf(v)))


def f32_table(vals):
    return np.array([[once_f32(v.real), once_f32(v.imag)] for v in vals], np.float32)


def f64_table(vals):
    return np.array([[float(v.real), float(v.imag)] for v in vals], np.float64)


def bits32(a):
    return np.asarray(a, np.float32).view(np.uint32)


def ulps64(a, b):
    """|a - b| in units of the last place of b (f64)."""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    sp = np.spacing(np.abs(b))
    sp[b == 0] = np.spacing(0.0)
    return np.abs(a - b) / sp


def device_twiddles(n, count, prec):
    import torch
    es = 8 if prec == FftPrecision.F32 else 16
    t = torch.empty(max(1, count) * es, dtype=torch.uint8, device="cuda")
    W.twiddle_table(t, n, count, prec)
    torch.cuda.synchronize()
    a = t.cpu().numpy().view(np.float32 if prec == FftPrecision.F32 else np.float64)
    return a.reshape(-1, 2)[:count]


def test_twiddle_golden_f32(torch_cuda):
    gold = json.load(open(os.path.join(GOLD, "twiddle_f32.json")))
    for n in sorted({g[0] for g in gold}):
        tab = bits32(device_twiddles(n, n, FftPrecision.F32))
        for gn, k, re, im in gold:
            if gn == n:
                assert (tab[k, 0], tab[k, 1]) == (re, im), (n, k, tab[k], re, im)


def test_twiddle_golden_f64(torch_cuda):
    gold = json.load(open(os.path.join(GOLD, "twiddle_f64.json")))
    worst = 0.0
    for n in sorted({g[0] for g in gold}):
        ks = [g[1] for g in gold if g[0] == n]
        tab = device_twiddles(n, max(ks) + 1, FftPrecision.F64)
        for gn, k, re, im in gold:
            if gn == n:
                want = [struct.unpack("<d", bytes.fromhex(h)[::-1])[0] for h in (re, im)]
                worst = max(worst, float(np.max(ulps64(tab[k], want))))
    print(f"f64 twiddles: worst {worst:.2f} ulp", flush=True)
    assert worst <= 1.0


@pytest.mark.parametrize("n", [4096, 1009, 3000])
def test_whole_f32_twiddle_table_is_rounded_once(torch_cuda, n):
    tab = device_twiddles(n, n, FftPrecision.F32)
    want = f32_table([exact(k, n) for k in range(n)])
    assert np.array_equal(bits32(tab), bits32(want))


def _stage_expected(radices, prec):
    vals = []
    L = 1
    for s, R in enumerate(radices):
        if s >= 1:
            for m in range(1, R):
                for k in range(L):
                    vals.append(exact(k * m, L * R))
        L *= R
    return vals


def _check_complex(desc, raw, vals, tag):
    if desc.get("prec", "f32") == "f32":
        got = raw.view(np.float32).reshape(-1, 2)[:len(vals)]
        want = f32_table(vals)
        bad = np.nonzero(bits32(got) != bits32(want))
        assert bad[0].size == 0, (tag, desc, got[bad[0][:4]], want[bad[0][:4]])
    else:
        got = raw.view(np.float64).reshape(-1, 2)[:len(vals)]
        u = float(np.max(ulps64(got, f64_table(vals)))) if len(vals) else 0.0
        assert u <= 1.0, (tag, desc, u)


def _filter_expected(desc):
    m = int(desc["m"])
    b = [mp.mpc(0)] * m
    if desc["kind"] == "bluestein-filter":
        n = int(desc["n"])
        for i in range(n):
            c = mp.conj(exact((i * i) % (2 * n), 2 * n))
            b[i] = c
            if i:
                b[m - i] = c
    else:
        p, h = int(desc["p"]), int(desc["ginv"])
        ln = p - 1
        for t in range(ln):
            b[t] = exact(pow(h, t, p), p)
        if m != ln:
            for idx in range(m - ln + 1, m):
                b[idx] = exact(pow(h, ln - (m - idx), p), p)
    seq = np.array([complex(v) for v in b], np.complex128)
    return np.fft.fft(seq) / m


PLANS = [
    ("c2c", [4096], 3, FftPrecision.F32, None),                  # register kernel stage twiddles
    ("c2c", [1000], 3, FftPrecision.F32, None),                  # runtime-schedule LDS kernel
    ("c2c", [256], 4, FftPrecision.F64, None),
    ("c2c", [127], 4, FftPrecision.F32, None),                   # Rader, cyclic M = 126
    ("c2c", [1009], 2, FftPrecision.F32, None),                  # Rader M = 1008
    ("c2c", [1009], 2, FftPrecision.F32, "bluestein"),           # forced Bluestein
    ("c2c", [3256], 2, FftPrecision.F32, None),                  # Bluestein (37 is not a radix)
    ("c2c", [10007], 1, FftPrecision.F32, "bluestein"),          # multi-pass Bluestein
    ("c2c", [2999], 1, FftPrecision.F64, None),                  # f64 Rader filter
    ("c2c", [47], 8, FftPrecision.F32, None),                    # direct prime roots
    ("c2c", [1 << 20], 1, FftPrecision.F32, None),               # split-pass twiddle factors
    ("c2c", [1 << 22], 1, FftPrecision.F32, None),
]


@pytest.mark.parametrize("i", range(len(PLANS)))
def test_plan_tables(torch_cuda, i):
    kind, shape, batch, prec, force = PLANS[i]
    cfg = FftConfig.new_nd(shape).with_batch(batch).with_precision(prec)
    if force == "bluestein":
        cfg = cfg.with_tuning(FftTuning().with_force_bluestein_axes([0]))
    plan = FftPlan.c2c(cfg)
    tabs = plan.tables()
    assert tabs, plan.describe()
    kinds = set()
    for desc, raw in tabs:
        k = desc["kind"]
        kinds.add(k)
        if k == "stage-twiddles":
            rad = [int(r) for r in desc["radices"].split(",")]
            vals = _stage_expected(rad, desc["prec"])
            assert len(vals) == int(desc["count"])
            _check_complex(desc, raw, vals, "stage")
        elif k == "chirp":
            n = int(desc["n"])
            _check_complex(desc, raw, [exact((j * j) % (2 * n), 2 * n) for j in range(n)], "chirp")
        elif k == "roots":
            n, c = int(desc["n"]), int(desc["count"])
            _check_complex(desc, raw, [exact(j, n) for j in range(c)], "roots")
        elif k == "roots-strided":
            n, c, s = int(desc["n"]), int(desc["count"]), int(desc["mult"])
            _check_complex(desc, raw, [exact((j * s) % n, n) for j in range(c)], "roots-strided")
        elif k in ("bluestein-filter", "rader-filter"):
            want = _filter_expected(desc)
            if desc["prec"] == "f32":
                got = raw.view(np.float32).reshape(-1, 2).astype(np.float64)
                w = np.stack([want.real, want.imag], axis=1)
                tol = np.spacing(np.abs(w).astype(np.float32)).astype(np.float64) + 1e-15 * np.max(np.abs(w))
                assert np.all(np.abs(got - w) <= tol), (desc, float(np.max(np.abs(got - w) / tol)))
            else:
                got = raw.view(np.complex128)
                assert np.max(np.abs(got - want)) <= 1e-14 * np.max(np.abs(want)), desc
        elif k in ("rader-perm-in", "rader-perm-out"):
            p = int(desc["p"])
            g = int(desc["g"] if k == "rader-perm-in" else desc["ginv"])
            got = raw.view(np.uint32)
            assert list(got) == [pow(g, q, p) for q in range(p - 1)], desc
            if k == "rader-perm-in":
                assert sorted(got) == list(range(1, p))  # a permutation of 1..p-1 (g generates)
        else:
            raise AssertionError(f"unknown table kind {k}")
    print(f"{shape} {prec.name} {force}: {sorted(kinds)} [{plan.describe()}]", flush=True)
