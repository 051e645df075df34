"""Writes tests/golden/*.json: pinning vectors for the FP64 oracle.

* kat.json       -- known-answer tests quoted from the wgpuFFT sources
                    (values, not code): src/math/mod.rs:575-700,
                    src/runtime/axis_policy.rs:185-266, src/runtime/mod.rs:113-126,
                    src/runtime/rader_axis.rs:2851-2856.
* dft_mp.json    -- DFTs of small seeded inputs evaluated with mpmath at 40
                    significant digits (independent of oracle/oracle.c), for
                    lengths that exercise every branch of the oracle recursion
                    (radix splits, direct primes <= 64, Bluestein primes > 64).
* twiddle_f32.json -- exp(-2 pi i k / N) at 40 digits rounded once to f32 bits,
                    for a few N and k (the "rounded once" contract of the
                    reference's canonical_twiddle_f32, src/runtime/twiddle.rs:198-204).
* twiddle_f64.json -- the same rounded once to f64 (bits as hex), for the
                    f64 tables (twiddle.rs:206-211).
* chirp_f32.json -- Bluestein chirps exp(-pi i (k^2 mod 2N) / N) rounded once to
                    f32 bits for N = 1009, 3256, 10007 (bluestein_axis.rs:818-867).

Run from the repository root:  python tests/golden/make_golden.py
"""
import json
import os
import struct

import mpmath as mp

HERE = os.path.dirname(os.path.abspath(__file__))
mp.mp.dps = 40


def splitmix64(seed):
    s = seed & 0xFFFFFFFFFFFFFFFF
    while True:
        s = (s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        yield z ^ (z >> 31)


def uniform(seed, count):
    g = splitmix64(seed)
    return [2.0 * ((next(g) >> 40) * 2.0 ** -24) - 1.0 for _ in range(count)]


def mp_dft(x, sign):
    n = len(x)
    out = []
    for k in range(n):
        acc = mp.mpc(0)
        for j in range(n):
            acc += x[j] * mp.expjpi(sign * mp.mpf(2 * ((j * k) % n)) / n)
        out.append(acc)
    return out


def f32_bits(v):
    """v rounded once (to nearest, ties to even) to f32, as its bit pattern."""
    with mp.workprec(24):
        r = +mp.mpf(v)
    return struct.unpack("<I", struct.pack("<f", float(r)))[0]


def f64_hex(v):
    """v rounded once to f64, as the hex of its bit pattern."""
    with mp.workprec(53):
        r = +mp.mpf(v)
    return "%016x" % struct.unpack("<Q", struct.pack("<d", float(r)))[0]


def main():
    kat = {
        "dft4_in": [1, 2, 3, 4],
        "dft4_out": [[10, 0], [-2, 2], [-2, 0], [-2, -2]],
        "r2c4_in": [1, 2, 3, 4],
        "r2c4_packed": [10, 0, -2, 2, -2, 0],
        "r2c_c2r_odd5": [0.25, -1.0, 2.0, 0.5, -0.75],
        "r2c_c2r_nd_3x2": [0.25, -1.0, 2.0, 0.5, -0.75, 1.25],
        "axis_kinds": {"8": "Mixed", "21": "Mixed", "17": "Rader", "29": "Rader", "34": "Bluestein"},
        "is_prime": {"0": False, "1": False, "2": True, "29": True, "34": False},
        "next_smooth": {"31": 32, "34": 35},
        "factor_acceptance": [2, 3, 4, 5, 6, 7, 8, 10, 11, 12, 13, 15, 16, 21],
        "factor_rejects": [17],
        "rader_perm_17_sorted": list(range(1, 17)),
        "uniform_seed1_first8": uniform(1, 8),
    }
    with open(os.path.join(HERE, "kat.json"), "w") as f:
        json.dump(kat, f, indent=1)

    cases = []
    for n in [1, 2, 3, 4, 5, 7, 12, 16, 17, 30, 49, 67, 97, 120, 128, 131]:
        v = uniform(500 + n, 2 * n)
        x = [mp.mpc(v[2 * i], v[2 * i + 1]) for i in range(n)]
        for sign, name in ((-1, "forward"), (1, "inverse")):
            y = mp_dft(x, sign)
            cases.append({"n": n, "seed": 500 + n, "direction": name,
                          "out": [[float(c.real), float(c.imag)] for c in y]})
    with open(os.path.join(HERE, "dft_mp.json"), "w") as f:
        json.dump(cases, f)

    tw = []
    for n in [7, 1009, 4096, 65536]:
        ks = sorted(set([0, 1, 2, n // 8, n // 4, n // 3, n // 2, n - 1] + [(k * 7919) % n for k in range(40)]))
        for k in ks:
            w = mp.expjpi(-mp.mpf(2 * k) / n)
            tw.append([n, k, f32_bits(w.real), f32_bits(w.imag)])
    with open(os.path.join(HERE, "twiddle_f32.json"), "w") as f:
        json.dump(tw, f)

    tw64 = []
    for n in [7, 1009, 4096, 65536, 1 << 24]:
        ks = sorted(set([0, 1, 2, n // 8, n // 4, n // 3, n // 2, n - 1] + [(k * 7919) % n for k in range(24)]))
        for k in ks:
            w = mp.expjpi(-mp.mpf(2 * k) / n)
            tw64.append([n, k, f64_hex(w.real), f64_hex(w.imag)])
    with open(os.path.join(HERE, "twiddle_f64.json"), "w") as f:
        json.dump(tw64, f)

    ch = []
    for n in [1009, 3256, 10007]:
        ks = sorted(set([0, 1, 2, 3, n // 2, n - 2, n - 1] + [(k * 7919) % n for k in range(40)]))
        for k in ks:
            w = mp.expjpi(-mp.mpf((k * k) % (2 * n)) / n)
            ch.append([n, k, f32_bits(w.real), f32_bits(w.imag)])
    with open(os.path.join(HERE, "chirp_f32.json"), "w") as f:
        json.dump(ch, f)


if __name__ == "__main__":
    main()
