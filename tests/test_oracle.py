"""Pins the FP64 oracle (oracle/oracle.c) before the tests trust it.

Each oracle function is checked against something it was not derived from:
the wgpuFFT known-answer tests (tests/golden/kat.json), mpmath DFTs at 40
digits (tests/golden/dft_mp.json), numpy's pocketfft, closed forms (impulse,
tone, constant, Parseval, linearity) and brute-force enumeration for the
integer helpers.  A planted mistake in any of them fails at least one test
here (see test_planted_mistakes_are_caught).
"""
import json
import math
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


KAT = load("kat.json")


# ---------------------------------------------------------------- KATs
def test_kat_length_four_dft():
    """src/math/mod.rs:575-592 (f32) and :594-613 (f64)."""
    y = O.c2c_nd(np.array(KAT["dft4_in"], dtype=complex), [4], normalization=O.NORM_NONE)
    exp = np.array([complex(*v) for v in KAT["dft4_out"]])
    assert np.max(np.abs(y - exp)) < 1e-12


def test_kat_r2c_packed_even_length():
    """src/math/mod.rs:683-689: [1,2,3,4] -> [10,0,-2,2,-2,0]."""
    p = O.r2c_nd(np.array(KAT["r2c4_in"], dtype=float), [4], normalization=O.NORM_NONE)
    got = np.stack([p.real, p.imag], axis=1).reshape(-1)
    assert np.max(np.abs(got - np.array(KAT["r2c4_packed"]))) < 1e-12


def test_kat_r2c_c2r_round_trip_odd_length():
    """src/math/mod.rs:690-700 (forward Normalization::None, inverse default)."""
    x = np.array(KAT["r2c_c2r_odd5"])
    p = O.r2c_nd(x, [5], normalization=O.NORM_NONE)
    assert p.size == 3
    back = O.c2r_nd(p, [5])
    assert np.max(np.abs(back - x)) < 1e-12


def test_kat_r2c_c2r_nd_packs_axis_zero():
    """src/math/mod.rs:702-712: shape [3,2] packs axis 0 to 2 -> 4 complex."""
    x = np.array(KAT["r2c_c2r_nd_3x2"])
    p = O.r2c_nd(x, [3, 2], normalization=O.NORM_NONE)
    assert p.size == 4
    back = O.c2r_nd(p, [3, 2])
    assert np.max(np.abs(back - x)) < 1e-12


def test_kat_batched_odd_axis_subset_closed_form():
    """src/math/mod.rs:657-680: shape [2,3], axes [1], batch 2, inputs 1..12."""
    x = np.arange(1, 13).astype(complex)
    y = O.c2c_nd(x, [2, 3], axes=[1], batch=2, normalization=O.NORM_NONE)
    root = math.sqrt(3.0) * 0.5
    exp = np.zeros(12, dtype=complex)
    for b in range(2):
        base = b * 6
        for xx in range(2):
            a, bb, c = x[base + xx].real, x[base + 2 + xx].real, x[base + 4 + xx].real
            exp[base + xx] = a + bb + c
            exp[base + 2 + xx] = complex(a - 0.5 * (bb + c), root * (c - bb))
            exp[base + 4 + xx] = complex(a - 0.5 * (bb + c), root * (bb - c))
    assert np.max(np.abs(y - exp)) < 1e-12


# -------------------------------------------------- mpmath golden vectors
@pytest.mark.parametrize("case", load("dft_mp.json"), ids=lambda c: f"{c['n']}-{c['direction']}")
def test_fft1d_matches_mpmath(case):
    n = case["n"]
    v = O.uniform_f32(case["seed"], 2 * n).astype(np.float64)
    x = v[0::2] + 1j * v[1::2]
    y = O.fft1d(x, inverse=case["direction"] == "inverse")
    exp = np.array([complex(a, b) for a, b in case["out"]])
    peak = max(1.0, np.max(np.abs(exp)))
    # the reference's own bound for fft_f64 against a plain DFT (math/mod.rs:615-655)
    assert np.max(np.abs(y - exp)) < 1e-12 * peak


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 12, 16, 17, 30, 49, 97, 128, 210, 282, 1060, 1523])
def test_fft1d_matches_brute_force_long_double(n):
    """The reference's f64_fft_matches_a_plain_dft lengths (src/math/mod.rs:615-655)."""
    i = np.arange(n)
    x = np.sin(i * 0.37) + 1j * np.cos(i * 0.11)
    for inv in (False, True):
        y = O.fft1d(x, inverse=inv)
        exp = O.dft_naive(x, inverse=inv)
        peak = max(1.0, np.max(np.abs(exp)))
        assert np.max(np.abs(y - exp)) < 1e-12 * peak


@pytest.mark.parametrize("n", [6, 8, 11, 13, 64, 65, 100, 127, 243, 1000, 1009, 2048, 4093, 4096, 10007, 65536])
def test_fft1d_matches_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert O.rel_l2(O.fft1d(x), np.fft.fft(x)) < 1e-14
    assert O.rel_l2(O.fft1d(x, inverse=True), np.fft.ifft(x) * n) < 1e-14


# ------------------------------------------------------------ ND driver
ND = [([8], None, 3), ([6, 5], None, 2), ([4, 6, 10], None, 1), ([4, 6, 10], [2, 0], 2),
      ([5, 7], [1], 3), ([3, 4, 5, 2], [1, 3], 1), ([16, 9], [0], 2)]


@pytest.mark.parametrize("shape,axes,batch", ND)
@pytest.mark.parametrize("inverse", [False, True])
@pytest.mark.parametrize("norm", [O.NORM_NONE, O.NORM_FORWARD, O.NORM_INVERSE, O.NORM_ORTHO])
def test_c2c_nd_matches_numpy(shape, axes, batch, inverse, norm):
    """Axis 0 is the fastest axis: the flat buffer is a numpy array of the reversed shape."""
    rng = np.random.default_rng(len(shape) * 7 + batch)
    cnt = int(np.prod(shape)) * batch
    x = rng.standard_normal(cnt) + 1j * rng.standard_normal(cnt)
    y = O.c2c_nd(x, shape, axes, batch, inverse, norm, nthreads=3)
    arr = x.reshape([batch] + shape[::-1])
    rank = len(shape)
    ax = list(range(rank)) if axes is None else axes
    np_axes = [rank - a for a in ax]  # batch is numpy axis 0
    f = np.fft.ifftn(arr, axes=np_axes) * np.prod([shape[a] for a in ax]) if inverse else np.fft.fftn(arr, axes=np_axes)
    total = float(np.prod(shape))  # the reference scales by prod(shape) (config.rs:262-286)
    s = {O.NORM_NONE: 1.0, O.NORM_FORWARD: (1.0 if inverse else 1.0 / total),
         O.NORM_INVERSE: (1.0 / total if inverse else 1.0), O.NORM_ORTHO: 1.0 / math.sqrt(total)}[norm]
    assert O.rel_l2(y, f.reshape(-1) * s) < 1e-14
    assert O.scale(shape, inverse, norm) == pytest.approx(s, rel=1e-15)


def test_nthreads_does_not_change_results():
    rng = np.random.default_rng(3)
    x = rng.standard_normal(64 * 48) + 1j * rng.standard_normal(64 * 48)
    a = O.c2c_nd(x, [64, 48], nthreads=1)
    b = O.c2c_nd(x, [64, 48], nthreads=7)
    assert np.array_equal(a, b)


# ------------------------------------------------------------ closed forms
@pytest.mark.parametrize("n", [5, 16, 17, 120, 1009])
def test_closed_forms(n):
    k = np.arange(n)
    imp = np.zeros(n, dtype=complex)
    imp[1] = 1.0  # examples/quickstart.rs:14-33: delta[n-1] -> exp(-2 pi i k / n)
    assert np.max(np.abs(O.fft1d(imp) - np.exp(-2j * np.pi * k / n))) < 1e-13
    tone = np.exp(2j * np.pi * 3 * k / n)  # src/cpu.rs:1040-1078: tone -> n * delta
    exp = np.zeros(n, dtype=complex)
    exp[3 % n] = n
    assert np.max(np.abs(O.fft1d(tone) - exp)) < 1e-10
    const = np.full(n, 2.5 + 0.5j)
    y = O.fft1d(const)
    assert abs(y[0] - n * (2.5 + 0.5j)) < 1e-10 and np.max(np.abs(y[1:])) < 1e-10
    rng = np.random.default_rng(n)
    a = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    fa, fb = O.fft1d(a), O.fft1d(b)
    assert np.max(np.abs(O.fft1d(2 * a - 3j * b) - (2 * fa - 3j * fb))) < 1e-10  # linearity
    assert np.sum(np.abs(fa) ** 2) == pytest.approx(n * np.sum(np.abs(a) ** 2), rel=1e-12)  # Parseval


# ----------------------------------------------------------------- real
@pytest.mark.parametrize("shape,batch", [([4], 1), ([5], 2), ([16, 3], 1), ([9, 4, 5], 2), ([2048], 1)])
def test_r2c_matches_numpy_rfftn(shape, batch):
    """numpy's rfftn halves its last axis, which is axis 0 of the reversed shape."""
    rng = np.random.default_rng(sum(shape))
    x = rng.standard_normal(int(np.prod(shape)) * batch)
    p = O.r2c_nd(x, shape, batch)
    arr = x.reshape([batch] + shape[::-1])
    exp = np.fft.rfftn(arr, axes=list(range(1, len(shape) + 1)))
    assert O.rel_l2(p, exp.reshape(-1)) < 1e-14


@pytest.mark.parametrize("shape,batch", [([4], 1), ([5], 2), ([16, 3], 1), ([9, 4, 5], 2), ([10, 6], 1)])
def test_c2r_hermitian_input_matches_numpy_irfftn(shape, batch):
    rng = np.random.default_rng(sum(shape) + 1)
    x = rng.standard_normal(int(np.prod(shape)) * batch)
    arr = x.reshape([batch] + shape[::-1])
    axes = list(range(1, len(shape) + 1))
    p = np.fft.rfftn(arr, axes=axes)
    back = O.c2r_nd(p.reshape(-1), shape, batch)
    exp = np.fft.irfftn(p, s=shape[::-1], axes=axes)
    assert np.max(np.abs(back - exp.reshape(-1))) < 1e-12


@pytest.mark.parametrize("shape", [[4], [5], [6, 3], [5, 4, 2], [8, 8]])
def test_c2r_non_hermitian_projection_is_idempotent(shape):
    """C2R keeps only the Hermitian part of its input: C2R(R2C(C2R(P))) * N = N * C2R(P)."""
    rng = np.random.default_rng(len(shape))
    pk = int(np.prod(O.packed_shape(shape)))
    p = rng.standard_normal(pk) + 1j * rng.standard_normal(pk)
    r = O.c2r_nd(p, shape)
    p2 = O.r2c_nd(r, shape, normalization=O.NORM_NONE)
    r2 = O.c2r_nd(p2, shape)
    assert np.max(np.abs(r2 - r)) < 1e-13


def test_c2r_length4_closed_form_for_non_hermitian_input():
    """N=4, packed [a, b, c]: bins 0 and 2 are self-conjugate (imaginary parts
    dropped) and bin 3 mirrors conj(b) (src/math/mod.rs:114-153)."""
    a, b, c = 1 + 2j, 0.5 - 1j, -2 + 3j
    r = O.c2r_nd(np.array([a, b, c]), [4])
    full = np.array([a.real, b, c.real, np.conj(b)])
    exp = np.array([sum(full[k] * np.exp(2j * np.pi * k * t / 4) for k in range(4)).real / 4 for t in range(4)])
    assert np.max(np.abs(r - exp)) < 1e-14


# ------------------------------------------------------- integer helpers
def test_axis_kind_kats():
    """src/runtime/axis_policy.rs:185-189, via the oracle's integer helpers."""
    for n, kind in KAT["axis_kinds"].items():
        n = int(n)
        if O.factor_supported(n) is not None:
            got = "Mixed"
        elif O.is_prime(n) and n <= 8192:
            got = "Rader"
        else:
            got = "Bluestein"
        assert got == kind, n
    for n, p in KAT["is_prime"].items():
        assert O.is_prime(int(n)) == p
    for n, m in KAT["next_smooth"].items():
        assert O.next_smooth(int(n)) == m


def test_factorization_kats_and_greedy_order():
    for n in KAT["factor_acceptance"]:
        f = O.factor_supported(n)
        assert f is not None and int(np.prod(f)) == n and all(r in (2, 3, 4, 5, 7, 8, 11, 13) for r in f)
    for n in KAT["factor_rejects"]:
        assert O.factor_supported(n) is None
    # greedy [13,11,8,7,5,4,3,2] (src/runtime/mod.rs:31): the values in the hot-path table
    assert O.factor_supported(1024) == [8, 8, 8, 2]
    assert O.factor_supported(4096) == [8, 8, 8, 8]
    assert O.factor_supported(120) == [8, 5, 3]
    assert O.factor_supported(1001) == [13, 11, 7]
    assert O.factor_supported(343) == [7, 7, 7]


def test_smooth_and_prime_helpers_by_brute_force():
    def smooth(n):
        for p in (2, 3, 5, 7, 11, 13):
            while n % p == 0:
                n //= p
        return n == 1

    for n in range(1, 3000):
        assert O.is_prime(n) == (n > 1 and all(n % d for d in range(2, int(n ** 0.5) + 1)))
        m = O.next_smooth(n)
        assert m >= n and smooth(m) and all(not smooth(k) for k in range(n, m))
        assert (O.factor_supported(n) is not None) == smooth(n)


@pytest.mark.parametrize("p", [3, 5, 7, 17, 127, 1009, 2003, 8191])
def test_primitive_root_and_rader_permutation(p):
    def order(c):
        k, v = 1, c % p
        while v != 1:
            v = v * c % p
            k += 1
        return k

    g = O.primitive_root(p)
    # the smallest generator of the multiplicative group (axis_policy.rs:134-153)
    assert order(g) == p - 1 and all(order(c) < p - 1 for c in range(2, g))
    perm = O.rader_perm(p)
    assert perm == [pow(g, k + 1, p) for k in range(p - 1)]
    assert sorted(perm) == list(range(1, p))  # rader_axis.rs:2851-2856 for p = 17


@pytest.mark.parametrize("n", [127, 1009, 1381, 2003, 4093, 8191])
def test_convolution_lengths(n):
    def smooth(m):
        for q in (2, 3, 5, 7, 11, 13):
            while m % q == 0:
                m //= q
        return m == 1

    r = O.rader_conv_len(n)
    if smooth(n - 1):
        assert r == n - 1
    else:
        assert r >= 2 * (n - 1) - 1 and smooth(r) and not any(smooth(k) for k in range(2 * n - 3, r))
    b = O.bluestein_conv_len(n)
    assert b >= 2 * n - 1 and smooth(b) and not any(smooth(k) for k in range(2 * n - 1, b))
    for i in (0, 1, n - 1, n // 2, 12345 % n):
        assert O.bluestein_phase_index(n, i) == (i * i) % (2 * n)
    assert O.bluestein_phase_index(2 ** 31 - 1, 2 ** 31 - 2) == ((2 ** 31 - 2) ** 2) % (2 * (2 ** 31 - 1))


def test_df64_split_join_closed_form():
    rng = np.random.default_rng(1)
    v = rng.standard_normal(1000) * 10.0 ** rng.integers(-5, 5, 1000)
    hl = O.df64_split(v)
    hi = hl[0::2]
    lo = hl[1::2]
    assert np.array_equal(hi, v.astype(np.float32))
    assert np.array_equal(lo, (v - hi.astype(np.float64)).astype(np.float32))
    back = O.df64_join(hl)
    assert np.max(np.abs(back - v) / np.abs(v)) < 2.0 ** -46


def test_splitmix64_generator_matches_published_algorithm():
    exp = KAT["uniform_seed1_first8"]
    assert np.array_equal(O.uniform_f32(1, 8), np.array(exp, dtype=np.float32))
    u = O.uniform_f32(7, 100000)
    assert u.min() >= -1.0 and u.max() < 1.0 and abs(float(u.mean())) < 0.01
    z = O.normal_f32(3, 100001)
    assert abs(float(z.mean())) < 0.02 and abs(float(z.std()) - 1.0) < 0.02


def test_planted_mistakes_are_caught():
    """A sign flip, a dropped term or a transposed axis in the oracle's output is
    caught by the independent references used above."""
    case = load("dft_mp.json")[20]
    n = case["n"]
    v = O.uniform_f32(case["seed"], 2 * n).astype(np.float64)
    x = v[0::2] + 1j * v[1::2]
    exp = np.array([complex(a, b) for a, b in case["out"]])
    good = O.fft1d(x, inverse=case["direction"] == "inverse")
    assert np.max(np.abs(good - exp)) < 1e-12 * max(1, np.max(np.abs(exp)))
    for bad in (np.conj(good), np.roll(good, 1), good - x[0]):
        assert np.max(np.abs(bad - exp)) > 1e-6
    y = O.c2c_nd(np.arange(12).astype(complex), [3, 4])
    yt = O.c2c_nd(np.arange(12).astype(complex), [4, 3])
    assert np.max(np.abs(y - yt)) > 1e-3
