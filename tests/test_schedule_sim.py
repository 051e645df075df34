"""CPU simulations of the kernels' index maps against numpy's FFT (the
schedule-simulation tests the reference keeps for its generated kernels:
register_fft.rs:1119-1228, axis_plan.rs:4246-4306).  Each restates, in numpy,
the addressing a HIP kernel uses, so that a wrong index, sign or twiddle
exponent in that map fails here without a GPU:

* Stockham stages with the stage-twiddle table layout (RegFft, lds_stage,
  k_stockham_pass): butterfly j reads j + m N/R, twiddles W_{LR}^{(j mod L) m}
  at table entry twoff(s) + (m-1) L + (j mod L), writes (j - j mod L) R +
  (j mod L) + m L;
* the two-pass four-step and the three-pass split with the line
  decomposition i = g % S1, j = (g / S1) % S2, o = g / (S1 S2) and the planner's
  strides (plan.cpp four_step / three_pass);
* the twiddle recurrence of a butterfly's outputs (PostSeq);
* the direct prime DFT's output pairs (k_direct_prime), and the medium-prime
  stages of Rader convolutions (lds_stage_prime: symmetric pairs, radices
  17..61, inside the Stockham schedules of 613 ... 6841);
* the packed R2C store and the Hermitian-extended C2R load;
* the XCD tile remap (a bijection for every grid size);
* the small-volume gather order (k_small_volume).
"""
import numpy as np
import pytest

rng = np.random.default_rng(7)


def crand(n):
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def W(N, e):
    # exact integer phase reduction before the exponential (large exponents
    # would otherwise carry the angle's own rounding)
    return np.exp(-2j * np.pi * (np.asarray(e, dtype=np.int64) % N) / N)


def stage_table(rad):
    """Stage-twiddle table: for s >= 1 a block of (R_s - 1) L_s entries,
    entry (m-1) L_s + k = W_{L_s R_s}^{k m}."""
    tab, L = [], rad[0]
    for R in rad[1:]:
        for m in range(1, R):
            for k in range(L):
                tab.append(W(L * R, k * m))
        L *= R
    return np.array(tab, dtype=complex)


def twoff(rad, s):
    o, L = 0, rad[0]
    for i in range(1, s):
        o += (rad[i] - 1) * L
        L *= rad[i]
    return o


def stockham(x, rad, tab):
    N = len(x)
    L = 1
    for s, R in enumerate(rad):
        nr = N // R
        y = np.empty_like(x)
        for j in range(nr):
            k = j % L
            w = np.array([x[j + m * nr] for m in range(R)])
            if L > 1:
                for m in range(1, R):
                    w[m] *= tab[twoff(rad, s) + (m - 1) * L + k]
            w = np.fft.fft(w)
            for m in range(R):
                y[(j - k) * R + k + m * L] = w[m]
        x = y
        L *= R
    return x


@pytest.mark.parametrize("rad", [[16, 16, 8], [8, 8, 8, 4], [15, 8], [7, 7, 7], [13, 11, 7], [2, 9, 7],
                                 [16, 4, 5, 5], [16, 16, 16], [3], [9, 9, 5, 5]])
def test_stockham_stages(rad):
    N = int(np.prod(rad))
    x = crand(N)
    y = stockham(x, rad, stage_table(rad))
    assert np.abs(y - np.fft.fft(x)).max() <= 1e-9 * np.abs(np.fft.fft(x)).max()
    assert len(stage_table(rad)) == N - rad[0]


def line_base(g, S1, S2, si, sj, so):
    i = g % S1
    r = g // S1
    j = r % S2
    o = r // S2
    return i * si + j * sj + o * so, j


def run_lines(src, dst, n, lines, S1, S2, lin, lout, post=None):
    """One line pass: line g of n points at lin = (si, sj, so, sn) -> lout,
    optional four-step twiddle W_P^{j k} on store (post = P)."""
    for g in range(lines):
        bi, j = line_base(g, S1, S2, *lin[:3])
        bo, _ = line_base(g, S1, S2, *lout[:3])
        v = np.fft.fft(src[bi + lin[3] * np.arange(n)])
        if post:
            v = v * W(post, j * np.arange(n))
        dst[bo + lout[3] * np.arange(n)] = v


@pytest.mark.parametrize("n1,n2,S,O", [(8, 16, 1, 3), (16, 8, 4, 2), (4, 32, 1, 1), (12, 10, 3, 2)])
def test_four_step_two_pass(n1, n2, S, O):
    """plan.cpp four_step: pass 1 transforms n1-point columns (points at stride
    S n2) with W_n^{q2 k1} on store; pass 2 transforms n2-point rows and stores
    transposed (k = k1 + n1 k2)."""
    n = n1 * n2
    x = crand(S * n * O)
    ws = np.zeros_like(x)
    out = np.zeros_like(x)
    run_lines(x, ws, n1, S * O * n2, S, n2, (1, S, S * n, S * n2), (1, S, S * n, S * n2), post=n)
    run_lines(ws, out, n2, S * O * n1, S, n1, (1, S * n2, S * n, S), (1, S, S * n, S * n1))
    ref = np.fft.fft(x.reshape(O, n, S), axis=1).reshape(-1)
    assert np.abs(out - ref).max() <= 1e-9 * np.abs(ref).max()


@pytest.mark.parametrize("n3,n1,n2,O", [(4, 8, 16, 3), (16, 4, 8, 1), (8, 8, 8, 2)])
def test_three_pass_split(n3, n1, n2, O):
    """plan.cpp three_pass with the planner's line sets and layouts."""
    n, m = n3 * n1 * n2, n1 * n2
    x = crand(n * O)
    ws = np.zeros_like(x)
    ws2 = np.zeros_like(x)
    out = np.zeros_like(x)
    run_lines(x, ws, n3, O * m, 1, m, (1, 1, n, m), (1, 1, n, m), post=n)
    run_lines(ws, ws2, n2, O * n3 * n1, 1, n1, (1, 1, m, n1), (1, 1, m, n1), post=m)
    run_lines(ws2, out, n1, O * n3 * n2, n3, n2, (m, n1, n, 1), (1, n3, n, n3 * n2))
    ref = np.fft.fft(x.reshape(O, n), axis=1).reshape(-1)
    assert np.abs(out - ref).max() <= 1e-9 * np.abs(ref).max()


@pytest.mark.parametrize("n0,n1,O,F", [(8, 8, 1, 2), (4, 16, 3, 2), (16, 6, 2, 2), (6, 10, 1, 2),
                                        (8, 16, 2, 4), (4, 8, 1, 4), (6, 12, 3, 4)])
def test_fold_2d(n0, n1, O, F):
    """plan.cpp fold_2d: pass 1 transforms row groups r + b Q (b < F, M = F Q)
    of axis 0 (group q at line_base(q) with S1 = 1, S2 = Q, row b at b Q n0
    elements on) and stores z_c = W_M^{r c} sum_b y[r + b Q] W_F^{b c} at row
    r + c Q; pass 2 runs Q-point lines over the F parts (S1 = n0 columns, S2 =
    F parts) and stores part c's bin k at row F k + c."""
    Q = n1 // F
    x = crand(n0 * n1 * O)
    ws = np.zeros_like(x)
    out = np.zeros_like(x)
    fold = Q * n0
    for q in range(Q * O):  # the kernel's lines = n1 O, group q = g // F
        b, r = line_base(q, 1, Q, 0, n0, n0 * n1)
        y = np.array([np.fft.fft(x[b + k * fold + np.arange(n0)]) for k in range(F)])
        z = np.fft.fft(y, axis=0)  # radix-F DFT across the rows
        for c in range(F):
            ws[b + c * fold + np.arange(n0)] = W(n1, r * c) * z[c]
    run_lines(ws, out, Q, n0 * F * O, n0, F, (1, Q * n0, n0 * n1, n0), (1, n0, n0 * n1, F * n0))
    ref = np.fft.fft2(x.reshape(O, n1, n0), axes=(1, 2)).reshape(-1)
    assert np.abs(out - ref).max() <= 1e-9 * np.abs(ref).max()


@pytest.mark.parametrize("n0,n1,O", [(8, 8, 1), (6, 16, 3), (16, 4, 2)])
def test_fold_2d_r2c(n0, n1, O):
    """plan.cpp fold_2d for R2C (real load, packed store): real rows r and
    r + H (pitch n0, fold_in = H n0) transform to packed rows of P0 = n0/2 + 1
    bins (pitch P0, fold_out = H P0) folded as in the complex case; pass 2
    runs H-point lines over the P0 packed columns of both halves."""
    F, H, P0 = 2, n1 // 2, n0 // 2 + 1
    x = rng.standard_normal(n0 * n1 * O)
    ws = np.zeros(P0 * n1 * O, dtype=complex)
    out = np.zeros_like(ws)
    for q in range(H * O):
        bi, r = line_base(q, 1, H, 0, n0, n0 * n1)
        bo, _ = line_base(q, 1, H, 0, P0, P0 * n1)
        ya = np.fft.fft(x[bi + np.arange(n0)])[:P0]
        yb = np.fft.fft(x[bi + H * n0 + np.arange(n0)])[:P0]
        ws[bo + np.arange(P0)] = ya + yb
        ws[bo + H * P0 + np.arange(P0)] = W(n1, r) * (ya - yb)
    run_lines(ws, out, H, P0 * F * O, P0, F, (1, H * P0, P0 * n1, P0), (1, P0, P0 * n1, F * P0))
    ref = np.fft.fft(np.fft.rfft(x.reshape(O, n1, n0), axis=2), axis=1).reshape(-1)
    assert np.abs(out - ref).max() <= 1e-9 * np.abs(ref).max()


@pytest.mark.parametrize("n0,n1,O", [(8, 8, 1), (6, 16, 3), (16, 4, 2)])
def test_unfold_2d_c2r(n0, n1, O):
    """plan.cpp unfold_2d_c2r: pass A runs H-point inverse transforms over
    the packed columns of the even and odd rows (row 2k + c of part c) into
    halves (row c H + n); pass B loads rows n and n + H of the halves and
    forms E_0 + W_M^{-n} E_1 and E_0 - W_M^{-n} E_1 per packed bin (the last
    radix-2 stage of axis 1), then the Hermitian-extended inverse of axis 0."""
    H, P0 = n1 // 2, n0 // 2 + 1
    xr = rng.standard_normal(n0 * n1 * O)
    X = np.fft.fft(np.fft.rfft(xr.reshape(O, n1, n0), axis=2), axis=1).reshape(-1)
    ws = np.zeros_like(X)
    for g in range(P0 * 2 * O):  # pass A (inverse = conj(DFT(conj)))
        bi, _ = line_base(g, P0, 2, 1, P0, P0 * n1)
        bo, _ = line_base(g, P0, 2, 1, H * P0, P0 * n1)
        ws[bo + P0 * np.arange(H)] = np.fft.ifft(X[bi + 2 * P0 * np.arange(H)]) * H
    out = np.zeros(n0 * n1 * O)
    for q in range(H * O):  # pass B
        bi, n = line_base(q, 1, H, 0, P0, P0 * n1)
        bo, _ = line_base(q, 1, H, 0, n0, n0 * n1)
        e0, e1 = ws[bi + np.arange(P0)], ws[bi + H * P0 + np.arange(P0)]
        t = np.conj(W(n1, n))
        for c, row in enumerate((e0 + t * e1, e0 - t * e1)):
            out[bo + c * H * n0 + np.arange(n0)] = np.fft.irfft(row, n0) * n0
    assert np.abs(out / (n0 * n1) - xr).max() <= 1e-9 * np.abs(xr).max()


@pytest.mark.parametrize("n", [4, 6, 34, 64, 3418])
def test_half_length_real(n):
    """kern_util.hip k_r2c_post / k_c2r_pre: R2C of N0 = 2h real points as an
    h-point complex DFT of z[n] = x[2n] + i x[2n+1] plus one twiddle pass; the
    C2R from a packed spectrum whose bins 0 and h carry imaginary parts (they
    are dropped) through the h-point inverse, equal to the unnormalized
    full-length C2R."""
    h = n // 2
    x = rng.standard_normal(n)
    Z = np.fft.fft(x[0::2] + 1j * x[1::2])
    k = np.arange(h + 1)
    a, b = Z[k % h], np.conj(Z[(h - k) % h])
    X = (a + b) / 2 + W(n, k) * (a - b) / 2j
    assert np.abs(X - np.fft.rfft(x)).max() <= 1e-9 * np.abs(X).max()
    P = rng.standard_normal(h + 1) + 1j * rng.standard_normal(h + 1)
    Q = P.copy()
    Q[0] = Q[0].real
    Q[h] = Q[h].real
    k = np.arange(h)
    xa, xb = Q[k], np.conj(Q[h - k])
    Zi = (xa + xb) + 1j * (xa - xb) * np.conj(W(n, k))
    z = np.fft.ifft(Zi) * h
    got = np.empty(n)
    got[0::2], got[1::2] = z.real, z.imag
    want = np.fft.irfft(P, n) * n
    assert np.abs(got - want).max() <= 1e-9 * np.abs(want).max()


def test_post_twiddle_recurrence():
    """PostSeq: W_P^{j (p0 + m L)} for m < R from W_P^{j p0} and the step
    W_P^{j L} by repeated multiplication, within ~R ulp of the direct values."""
    P, L, R = 1 << 24, 4096, 16
    for j, p0 in [(1, 0), (65535, 4095), (12345, 17), (65535, 1)]:
        w, step = W(P, j * p0), W(P, j * L)
        for m in range(R):
            assert abs(w - W(P, j * (p0 + m * L))) < 4 * R * 2.2e-16
            w = w * step


def prime_butterfly(x):
    """lds_stage_prime's medium-prime DFT: inputs m and R - m as their sum and
    difference, X_0 = x_0 + sum ap, X_q = A - iB and X_{R-q} = A + iB with
    A = x_0 + sum_m cos(2 pi mq/R) ap_m, B = sum_m sin(2 pi mq/R) am_m."""
    R = len(x)
    h = (R - 1) // 2
    ap = [x[m] + x[R - m] for m in range(1, h + 1)]
    am = [x[m] - x[R - m] for m in range(1, h + 1)]
    y = np.empty(R, complex)
    y[0] = x[0] + sum(ap)
    for q in range(1, h + 1):
        A, B = x[0], 0
        for m in range(1, h + 1):
            e = (m * q) % R
            A = A + np.cos(2 * np.pi * e / R) * ap[m - 1]
            B = B + np.sin(2 * np.pi * e / R) * am[m - 1]
        y[q] = A - 1j * B
        y[R - q] = A + 1j * B
    return y


# Schedules of the medium-prime Rader convolutions (p - 1 points, the medium
# primes first): 613, 1381, 1999, 4093, 4241, 5003, 6841
@pytest.mark.parametrize("rad", [[17, 4, 9], [23, 4, 15], [37, 6, 9], [31, 4, 3, 11], [53, 16, 5], [61, 41, 2],
                                 [19, 8, 9, 5]])
def test_medium_prime_stages(rad):
    N = int(np.prod(rad))
    x0 = crand(N)
    x = x0.copy()
    tab = stage_table(rad)
    L = 1
    for s, R in enumerate(rad):
        nr = N // R
        y = np.empty_like(x)
        for j in range(nr):
            k = j % L
            w = np.array([x[j + m * nr] for m in range(R)])
            if L > 1:
                for m in range(1, R):
                    w[m] *= tab[twoff(rad, s) + (m - 1) * L + k]
            w = prime_butterfly(w) if R > 16 else np.fft.fft(w)
            for m in range(R):
                y[(j - k) * R + k + m * L] = w[m]
        x, L = y, L * R
    ref = np.fft.fft(x0)
    assert np.abs(x - ref).max() <= 1e-9 * np.abs(ref).max()


@pytest.mark.parametrize("R", [17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61])
def test_prime_butterfly(R):
    x = crand(R)
    assert np.abs(prime_butterfly(x) - np.fft.fft(x)).max() <= 1e-9 * np.abs(np.fft.fft(x)).max()


@pytest.mark.parametrize("p", [47, 59, 83, 103, 107])
def test_direct_prime_pairs(p):
    """k_direct_prime: X_k = x_0 + sum s_n cos - i sum d_n sin and X_{p-k} with
    the opposite sign, s_n = x_n + x_{p-n}, d_n = x_n - x_{p-n}."""
    x = crand(p)
    X = np.zeros(p, complex)
    X[0] = x.sum()
    h = (p - 1) // 2
    for k in range(1, h + 1):
        A = B = 0
        t = 0
        for n in range(1, h + 1):
            t = (t + k) % p
            w = W(p, t)
            c, sn = w.real, -w.imag
            A += (x[n] + x[p - n]) * c
            B += (x[n] - x[p - n]) * sn
        X[k] = x[0] + A - 1j * B
        X[p - k] = x[0] + A + 1j * B
    assert np.abs(X - np.fft.fft(x)).max() <= 1e-9 * np.abs(X).max()


@pytest.mark.parametrize("n", [16, 15, 2048])
def test_real_modes(n):
    """LD_REAL + ST_PACKED stores bins k <= n/2; LD_HERMITIAN extends the packed
    half (bin p > n/2 = conj of bin n - p; the imaginary part of bins 0 and n/2
    dropped)."""
    x = rng.standard_normal(n)
    full = np.fft.fft(x)
    packed = full[: n // 2 + 1]
    assert np.allclose(packed, np.fft.rfft(x))
    ext = np.empty(n, complex)
    for p in range(n):
        up = p > n // 2
        v = packed[n - p if up else p]
        im = 0.0 if (p == 0 or 2 * p == n) else (-v.imag if up else v.imag)
        ext[p] = v.real + 1j * im
    assert np.allclose(np.fft.ifft(ext).imag, 0, atol=1e-12)
    assert np.allclose(np.fft.ifft(ext).real, x)


def test_xcd_remap_bijective():
    """xcd_remap: XCD x = b mod 8 holds q + (x < r) blocks (n = 8q + r), given
    the contiguous tiles x q + min(x, r) + b / 8."""
    for n in list(range(1, 300)) + [1025, 4096, 65537]:
        q, r = n >> 3, n & 7
        tiles = sorted((b & 7) * q + min(b & 7, r) + (b >> 3) for b in range(n))
        assert tiles == list(range(n))


@pytest.mark.parametrize("dims", [[16, 16], [12, 10, 9], [4, 256], [2, 2, 2, 2, 2, 2, 2, 2]])
def test_small_volume_gather(dims):
    """k_small_volume: per axis a (length n, stride S = prod(dims[:a])), thread
    t takes line l and point p (p fastest on axis 0, l fastest otherwise), line
    l = (i, o) with i = l % S, point at o S n + i + p S.  Every element is
    gathered exactly once and the axis FFTs give the full DFT."""
    vol = crand(int(np.prod(dims)))
    x0 = vol.copy()
    e = vol.size
    S = 1
    for n in dims:
        lines = e // n
        buf = np.empty((lines, n), complex)
        seen = np.zeros(e, int)
        for t in range(e):
            if S == 1:
                l, p = t // n, t % n
            else:
                p, l = t // lines, t % lines
            o, i = l // S, l % S
            idx = o * S * n + i + p * S
            buf[l, p] = vol[idx]
            seen[idx] += 1
        assert (seen == 1).all()
        res = np.fft.fft(buf, axis=1)
        for t in range(e):
            l, p = t // n, t % n
            o, i = l // S, l % S
            vol[o * S * n + i + p * S] = res[l, p]
        S *= n
    # axis 0 is the fastest-varying index: numpy's C order reversed
    ref = np.fft.fftn(x0.reshape(dims[::-1])).reshape(-1)
    assert np.abs(vol - ref).max() <= 1e-9 * np.abs(ref).max()


def jof(s, t, q, ns, nb, tpl, adj):
    """Butterfly q of thread t in stage s (RegFft::jof): adjacent pairs in the
    HBM stages (0 and the last) when adj, else strided by the thread count."""
    return t * nb + q if adj and (s == 0 or s == ns - 1) else t + q * tpl


@pytest.mark.parametrize("rad,tpl", [([8, 8, 8, 8], 256), ([8, 8, 8, 8], 128), ([16, 16, 16], 128),
                                     ([8, 8, 4, 4], 64), ([16, 16, 4], 32), ([8, 8, 8, 4], 128)])
@pytest.mark.parametrize("adj", [False, True])
def test_register_kernel_thread_map(rad, tpl, adj):
    """RegFft's per-thread schedule: stage-0 loads into register slots
    (q, m) of butterfly jof(0, t, q), LDS exchange writes with the computing
    stage's map and reads with the next stage's, last-stage stores from slot
    (q, m) to output j + m L.  With adj, slots (q, m) and (q + 1, m) load
    and store adjacent elements (one 16-byte access)."""
    N = int(np.prod(rad))
    ns = len(rad)
    x = crand(N)
    tab = stage_table(rad)
    regs = {}
    nr = N // rad[0]
    nb = -(-nr // tpl)
    for t in range(tpl):
        for q in range(nb):
            j = jof(0, t, q, ns, nb, tpl, adj)
            for m in range(rad[0]):
                regs[(t, q, m)] = x[j + m * nr]
            if adj and q % 2 == 0:
                # the pair load covers elements j + m nr and j + 1 + m nr
                assert jof(0, t, q + 1, ns, nb, tpl, adj) == j + 1
    out = np.zeros(N, dtype=complex)
    L = 1
    for s, R in enumerate(rad):
        nr = N // R
        nb = -(-nr // tpl)
        for t in range(tpl):
            for q in range(nb):
                j = jof(s, t, q, ns, nb, tpl, adj)
                if j >= nr:
                    continue
                k = j % L
                w = np.array([regs[(t, q, m)] for m in range(R)])
                if L > 1:
                    for m in range(1, R):
                        w[m] *= tab[twoff(rad, s) + (m - 1) * L + k]
                w = np.fft.fft(w)
                for m in range(R):
                    regs[(t, q, m)] = w[m]
        if s == ns - 1:
            for t in range(tpl):
                for q in range(nb):
                    j = jof(s, t, q, ns, nb, tpl, adj)
                    if j >= nr:
                        continue
                    k = j % L
                    for m in range(R):
                        out[(j - k) * R + k + m * L] = regs[(t, q, m)]
                    if adj and q % 2 == 0:
                        j1 = jof(s, t, q + 1, ns, nb, tpl, adj)
                        assert (j1 - j1 % L) * R + j1 % L == (j - k) * R + k + 1
            break
        # exchange: write with stage s's map, read with stage s + 1's
        lds = np.zeros(N, dtype=complex)
        for t in range(tpl):
            for q in range(nb):
                j = jof(s, t, q, ns, nb, tpl, adj)
                if j >= nr:
                    continue
                k = j % L
                for m in range(R):
                    lds[(j - k) * R + k + m * L] = regs[(t, q, m)]
        L *= R
        R2 = rad[s + 1]
        nr2 = N // R2
        nb2 = -(-nr2 // tpl)
        regs = {}
        for t in range(tpl):
            for q in range(nb2):
                j = jof(s + 1, t, q, ns, nb2, tpl, adj)
                if j >= nr2:
                    continue
                for m in range(R2):
                    regs[(t, q, m)] = lds[j + m * nr2]
    np.testing.assert_allclose(out, np.fft.fft(x), rtol=1e-9, atol=1e-9 * N)


def _reg_fft_threads(w, rad, tpl, tab):
    """RegBluestein::stages: registers w[t][q R_0 + m] hold point j + m M/R_0
    (j = t + q tpl); stage s multiplies register q R + m by table entry
    twoff(s) + (m-1) L + (j mod L), runs the radix-R DFT, and exchanges
    through LDS index (j - j mod L) R + (j mod L) + m L, read back as point
    j + m M/R_{s+1}.  Returns the registers after the last stage."""
    M = int(np.prod(rad))
    L = 1
    for s, R in enumerate(rad):
        nr = M // R
        nb = nr // tpl
        for t in range(tpl):
            for q in range(nb):
                j = t + q * tpl
                k = j % L
                v = np.array([w[t][q * R + m] for m in range(R)])
                if L > 1:
                    for m in range(1, R):
                        v[m] *= tab[twoff(rad, s) + (m - 1) * L + k]
                v = np.fft.fft(v)
                for m in range(R):
                    w[t][q * R + m] = v[m]
        if s + 1 == len(rad):
            break
        lds = np.zeros(M, dtype=complex)
        for t in range(tpl):
            for q in range(nb):
                j = t + q * tpl
                k = j % L
                for m in range(R):
                    lds[(j - k) * R + k + m * L] = w[t][q * R + m]
        L *= R
        R2 = rad[s + 1]
        nr2 = M // R2
        for t in range(tpl):
            w[t] = [0] * len(w[t])
            for q in range(nr2 // tpl):
                j = t + q * tpl
                for m in range(R2):
                    w[t][q * R2 + m] = lds[j + m * nr2]
    return w


@pytest.mark.parametrize("n,rad,tpl", [(37, [8, 2, 8], 8), (100, [8, 8, 8], 32), (509, [16, 4, 16], 64),
                                       (1000, [16, 8, 16], 128), (31, [8, 8], 8)])
def test_register_bluestein_map(n, rad, tpl):
    """RegBluestein::run: a[p] = x[p] c[p] loaded into the first stage's
    registers, forward transform, register q R_0 + m multiplied by F[p] and
    swapped with p = j + m M/R_0 (the last stage's output index of that
    register, since R_0 = R_last), forward transform again, X[p] = swap(.)
    c[p] for p < n: the DFT of x."""
    M = int(np.prod(rad))
    assert rad[0] == rad[-1] and M >= 2 * n - 1
    x = crand(n)
    c = W(2 * n, np.arange(n, dtype=np.int64) ** 2)          # chirp W_2n^(p^2)
    seq = np.zeros(M, dtype=complex)
    seq[:n] = np.conj(c)
    seq[M - n + 1:] = np.conj(c[1:][::-1])
    F = np.fft.fft(seq) / M
    tab = stage_table(rad)
    R0 = rad[0]
    nr0 = M // R0
    nb0 = nr0 // tpl
    w = [[0j] * (M // tpl) for _ in range(tpl)]
    for t in range(tpl):
        for q in range(nb0):
            j = t + q * tpl
            for m in range(R0):
                p = j + m * nr0
                w[t][q * R0 + m] = x[p] * c[p] if p < n else 0
    w = _reg_fft_threads(w, rad, tpl, tab)
    for t in range(tpl):
        for q in range(nb0):
            j = t + q * tpl
            for m in range(R0):
                p = j + m * nr0
                v = w[t][q * R0 + m] * F[p]
                w[t][q * R0 + m] = complex(v.imag, v.real)
    w = _reg_fft_threads(w, rad, tpl, tab)
    out = np.zeros(n, dtype=complex)
    for t in range(tpl):
        for q in range(nb0):
            j = t + q * tpl
            for m in range(R0):
                p = j + m * nr0
                if p < n:
                    v = w[t][q * R0 + m]
                    out[p] = complex(v.imag, v.real) * c[p]
    np.testing.assert_allclose(out, np.fft.fft(x), rtol=1e-9, atol=1e-9 * n)
