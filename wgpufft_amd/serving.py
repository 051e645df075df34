"""Continuous-batching FFT serving on one GPU (mifft_server_* in the C ABI).

``FftServer.replay`` runs an arrival trace through the server's event loop
(serve.cpp) and returns per-request arrival / admission / completion times;
``request_mix`` draws the seeded request population of the serving workload:

* 1D f32 and f64 C2C, lengths log-uniform in [64, 2^20] and, with equal
  probability, rounded to a {2,3,5,7,11,13}-smooth length, to a prime (Rader up
  to 8192, Bluestein above) or to a composite that is neither (Bluestein);
* 2D f32 C2C with sides log-uniform in [16, 1024] (smooth);
* R2C f32 on 1D and 2D real shapes of the same ranges;
each with uniform[-1,1) data generated on the device from SplitMix64 seeds.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import FftDirection, FftError, Normalization, _buf, _check, _lib, fill_uniform_f32

C2C, R2C = 0, 1
F32, F64 = 0, 1


def _smooth(n):
    for p in (2, 3, 5, 7, 11, 13):
        while n % p == 0:
            n //= p
    return n == 1


def _is_prime(n):
    if n < 2:
        return False
    if n % 2 == 0:
        return n == 2
    d = 3
    while d * d <= n:
        if n % d == 0:
            return False
        d += 2
    return True


def _next(n, pred):
    while not pred(n):
        n += 1
    return n


def length_menu(lo=64, hi=2 ** 20, points=16):
    """Log-spaced anchors and their smooth / prime / neither variants."""
    out = []
    for k in range(points):
        a = int(round(lo * (hi / lo) ** (k / (points - 1))))
        sm = _next(a, _smooth)
        if sm > hi:
            sm = max(m for m in range(a, 0, -1) if _smooth(m))
        pr = _next(a, _is_prime)
        bl = _next(a, lambda m: not _smooth(m) and not _is_prime(m))
        out.append((min(sm, hi), min(pr, hi + 1000), bl))
    return out


def request_mix(count, seed=6):
    """Seeded list of request specs: dicts kind, precision, shape, batch."""
    rng = np.random.default_rng(seed)
    menu = length_menu()
    sides = [16, 32, 48, 64, 96, 128, 192, 256, 384, 512, 768, 1024]
    reqs = []
    for _ in range(count):
        c = rng.random()
        if c < 0.4 or c < 0.6:
            prec = F32 if c < 0.4 else F64
            anchor = menu[int(rng.integers(len(menu)))]
            n = anchor[int(rng.integers(3))]
            reqs.append(dict(kind=C2C, precision=prec, shape=[n], batch=1))
        elif c < 0.8:
            reqs.append(dict(kind=C2C, precision=F32,
                             shape=[sides[int(rng.integers(len(sides)))], sides[int(rng.integers(len(sides)))]],
                             batch=1))
        else:
            if rng.random() < 0.5:
                # real axis 0 must fit one workgroup (<= 8192 points)
                small = [m[0] for m in menu if m[0] <= 8192]
                n = small[int(rng.integers(len(small)))]
                reqs.append(dict(kind=R2C, precision=F32, shape=[n], batch=1))
            else:
                reqs.append(dict(kind=R2C, precision=F32,
                                 shape=[sides[int(rng.integers(len(sides)))], sides[int(rng.integers(len(sides)))]],
                                 batch=1))
    return reqs


def poisson_arrivals(count, rate_hz, seed=6):
    rng = np.random.default_rng(seed + 1)
    return np.cumsum(rng.exponential(1.0 / rate_hz, count))


def request_bytes(spec):
    n = int(np.prod(spec["shape"])) * spec["batch"]
    es = 8 if spec["precision"] == F32 else 16
    if spec["kind"] == R2C:
        packed = (spec["shape"][0] // 2 + 1) * int(np.prod(spec["shape"][1:])) * spec["batch"]
        return 4 * n, 8 * packed
    return es * n, es * n


def allocate(specs, seed0=1000):
    """Device input/output buffers for each spec, inputs filled from SplitMix64."""
    import torch
    bufs = []
    for k, s in enumerate(specs):
        ib, ob = request_bytes(s)
        x = torch.empty(ib, dtype=torch.uint8, device="cuda")
        fill_uniform_f32(x, seed0 + k)
        y = torch.empty(ob, dtype=torch.uint8, device="cuda")
        bufs.append((x, y))
    return bufs


class FftServer:
    def __init__(self, budget_bytes: int = 64 << 30, streams: int = 4, max_tick_requests: int = 256,
                 batch_max_points: int | None = None):
        """batch_max_points: requests of at least this many points run through
        their resident plans beside the batched launch (None keeps the
        library default, 2^19)."""
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        _check(self._lib.mifft_server_create(int(budget_bytes), int(streams), int(max_tick_requests),
                                             ctypes.byref(h)))
        self._h = h
        if batch_max_points is not None:
            _check(self._lib.mifft_server_set_batch_limit(h, int(batch_max_points)))

    @staticmethod
    def _requests(specs, bufs):
        arr = (_lib.ServerRequest * len(specs))()
        for k, (s, (x, y)) in enumerate(zip(specs, bufs)):
            arr[k].kind = s["kind"]
            arr[k].precision = s["precision"]
            arr[k].rank = len(s["shape"])
            arr[k].direction = int(s.get("direction", FftDirection.Forward))
            arr[k].normalization = int(s.get("normalization", Normalization.Inverse))
            for i, v in enumerate(s["shape"]):
                arr[k].shape[i] = v
            arr[k].batch = s["batch"]
            ip, ib = _buf(x)
            op, ob = _buf(y)
            arr[k].in_ptr, arr[k].in_bytes = ip, ib
            arr[k].out_ptr, arr[k].out_bytes = op, ob
        return arr

    def prepare(self, specs, bufs):
        arr = self._requests(specs, bufs)
        _check(self._lib.mifft_server_prepare(self._h, arr, len(specs)))

    def replay(self, specs, bufs, arrivals, mode=0):
        """mode 0 continuous batching, 1 sequential per-plan dispatch; returns a
        numpy record array (arrival_s, admit_s, done_s, status, path)."""
        arr = self._requests(specs, bufs)
        t = (ctypes.c_double * len(specs))(*[float(a) for a in arrivals])
        res = (_lib.ServerResult * len(specs))()
        _check(self._lib.mifft_server_replay(self._h, arr, t, len(specs), int(mode), res))
        out = np.zeros(len(specs), dtype=[("arrival_s", "f8"), ("admit_s", "f8"), ("done_s", "f8"),
                                          ("status", "i4"), ("path", "i4")])
        for k in range(len(specs)):
            out[k] = (res[k].arrival_s, res[k].admit_s, res[k].done_s, res[k].status, res[k].path)
        return out

    # online serving (mifft_server_submit / _poll)
    def submit(self, spec, buffers) -> int:
        """Submits one request (a spec as from request_mix and its (input,
        output) buffers, kept alive by the caller until completion); returns
        its ticket."""
        arr = self._requests([spec], [buffers])
        t = ctypes.c_uint64()
        _check(self._lib.mifft_server_submit(self._h, arr, ctypes.byref(t)))
        return t.value

    def poll(self, wait: bool = False, cap: int = 1024):
        """Completions since the last poll: a list of dicts ticket, submit_s,
        admit_s, done_s, status, path."""
        out = (_lib.ServerCompletion * cap)()
        n = ctypes.c_int32()
        _check(self._lib.mifft_server_poll(self._h, int(bool(wait)), out, cap, ctypes.byref(n)))
        return [dict(ticket=out[k].ticket, submit_s=out[k].submit_s, admit_s=out[k].admit_s,
                     done_s=out[k].done_s, status=out[k].status, path=out[k].path) for k in range(n.value)]

    def drain(self):
        """Polls until nothing is queued or in flight; returns every completion."""
        got = []
        while True:
            got += self.poll(wait=True)
            q, f, c = self.pending()
            if q == 0 and f == 0 and c == 0:
                return got

    def pending(self):
        q, f, c = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        _check(self._lib.mifft_server_pending(self._h, ctypes.byref(q), ctypes.byref(f), ctypes.byref(c)))
        return q.value, f.value, c.value

    def set_plan_cache(self, max_shapes: int) -> None:
        _check(self._lib.mifft_server_set_plan_cache(self._h, int(max_shapes)))

    def cached_shapes(self) -> int:
        return int(self._lib.mifft_server_cached_shapes(self._h))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.mifft_server_destroy(h)
            self._h = None


def latency_summary(res):
    ok = res[res["status"] == 0]
    lat = (ok["done_s"] - ok["arrival_s"]) * 1e3
    span = float(res["done_s"].max() - res["arrival_s"].min())
    return {"requests": int(len(res)), "completed": int(len(ok)), "rejected": int((res["status"] != 0).sum()),
            "p50_ms": float(np.percentile(lat, 50)) if len(lat) else math.nan,
            "p99_ms": float(np.percentile(lat, 99)) if len(lat) else math.nan,
            "mean_ms": float(lat.mean()) if len(lat) else math.nan,
            "throughput_rps": float(len(ok) / span) if span > 0 else math.nan,
            "makespan_s": span,
            "batched_fraction": float((ok["path"] == 0).mean()) if len(ok) else math.nan}
