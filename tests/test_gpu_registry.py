"""Every compiled register-kernel variant (the autotuner's search space,
mifft_registry_name) forced through the tuning table onto a plan pass of its
shape, checked against the FP64 oracle in both directions.

A pool of plans (1D contiguous, strided single axes, R2C / C2R, two- and
three-pass splits of long axes, 2D folds, f64 / df64) is built once; every
launch reports its tuning key (mifft_plan_launch_tuning).  For each variant
the smallest pool plan with a launch of the variant's shape is rebuilt with
`key = variant`, the launch must report the variant, and the output must match
the oracle: f32 relative L2 <= 1e-6 (tests/gpu_accuracy.rs:94-104), f64
<= 1e-13 (tests/gpu_f64.rs:15), df64 <= 1e-11 (tests/gpu_df64.rs:14).
Persistent pipelined variants get more tiles than their grid has workgroups,
so that each workgroup loops."""
import re

import numpy as np
import pytest

import oracle
from tests.helpers import c2c_expected, empty_dev, from_dev, seeded_complex, to_dev
import wgpufft_amd as W
from wgpufft_amd import FftConfig, FftPlan, FftPrecision

pytestmark = pytest.mark.gpu

try:
    NAMES = W.registry_names()
except Exception:  # library not built: collection still succeeds, the GPU run fails loudly
    NAMES = []

PREC = {"f32": FftPrecision.F32, "f64": FftPrecision.F64, "df64": FftPrecision.Df64}
TOL = {"f32": 1e-6, "f64": 1e-13, "df64": 1e-11}
NAME_RE = re.compile(r"reg:(\w+):n(\d+):t(\d+):l(\d+):lf(\d)(\d):m(\d)/(\d):p(\d):v(\d):([0-9x]+)(:pipe\d)?(:fold)?(:adj)?$")


def parse_name(name):
    m = NAME_RE.match(name)
    assert m, name
    prec, n, tpl, lines, a, b, lm, sm, post, vec, rad, pipe, fold, adj = m.groups()
    return dict(prec=prec, n=int(n), tpl=int(tpl), lines=int(lines), lf=a + b, m=f"{lm}/{sm}", post=int(post),
                vec=int(vec), rad=rad, pipe=bool(pipe), fold=bool(fold), adj=bool(adj))


def key_matches(v, key):
    """Whether tuning key `key` (prec|nN|lfAB|mL/S|pP|pairQ|lines2^K[|foldF]) takes variant v."""
    parts = key.split("|")
    if len(parts) < 7:
        return False
    prec, n, lf, mm, post, pair = parts[0], parts[1], parts[2], parts[3], parts[4], parts[5]
    fold = parts[7] if len(parts) > 7 else ""
    if prec != v["prec"] or n != f"n{v['n']}" or lf != f"lf{v['lf']}" or mm != f"m{v['m']}":
        return False
    if post != f"p{v['post']}":
        return False
    if v["fold"]:
        return fold == f"fold{v['vec']}"
    if fold:
        return False
    return v["vec"] == 1 or pair == "pair1"


# The pool: (kind, shape, axes, batch, precision, extra tuning text).  Line
# counts are odd where a one-line-per-thread strided variant must apply.
def _pool():
    P = []
    for n in (64, 128, 256, 512, 1024, 2048, 4096):
        P.append(("c2c", [n], None, 37, "f32", ""))           # contiguous lines
        P.append(("c2c", [n], None, 5000, "f32", ""))         # more tiles than a persistent grid
        P.append(("c2c", [40, n], [1], 1, "f32", ""))         # strided, adjacent line pairs
        P.append(("c2c", [37, n], [1], 1, "f32", ""))         # strided, odd line count
        P.append(("r2c", [n], None, 37, "f32", ""))
        P.append(("c2r", [n], None, 37, "f32", ""))
        P.append(("r2c", [n], None, 1500, "f32", ""))
        P.append(("c2r", [n], None, 1500, "f32", ""))
    P.append(("c2c", [4400, 2048], [1], 1, "f32", ""))        # strided pipelined columns
    for n0 in (1024, 2048, 4096):
        P.append(("c2c", [n0, 512], None, 1, "f32", ""))      # 2D fold (F = 2)
        P.append(("c2c", [n0, 512], None, 1, "f32", f"fold|f32|n{n0}x512 = 4\n"))
        P.append(("c2c", [n0, 2048], None, 1, "f32", ""))     # fold rows: more tiles than a grid
        P.append(("c2c", [n0, 4096], None, 1, "f32", ""))
        P.append(("r2c", [n0, 512], None, 1, "f32", ""))
        P.append(("c2r", [n0, 512], None, 1, "f32", ""))
    for lg in range(12, 25):
        for split in ("2", "3"):
            b = 2 if lg <= 20 else 1
            P.append(("c2c", [1 << lg], None, b, "f32", f"split|f32|n{1 << lg} = {split}\n"))
    # split passes of a strided long axis (two adjacent columns: pairable)
    for lg in range(12, 25):
        for split in ("2", "3"):
            P.append(("c2c", [2, 1 << lg], [1], 1, "f32", f"split|f32|n{1 << lg} = {split}\n"))
    for n in (120, 343, 1001):
        P.append(("c2c", [n], None, 37, "f32", ""))
    for prec in ("f64", "df64"):
        for lg in (20, 22, 24):
            P.append(("c2c", [1 << lg], None, 1, prec, f"split|{prec}|n{1 << lg} = 2\n"))
        P.append(("c2c", [256], None, 37, prec, ""))
        P.append(("c2c", [37, 256], [1], 1, prec, ""))
        P.append(("c2c", [40, 256], [1], 1, prec, ""))
        P.append(("c2c", [4096], None, 5, prec, ""))
        P.append(("c2c", [37, 4096], [1], 1, prec, ""))
        P.append(("c2c", [1 << 16], None, 2, prec, ""))
        P.append(("c2c", [1 << 20], None, 1, prec, ""))
        P.append(("c2c", [1 << 24], None, 1, prec, ""))
    return P


POOL = _pool()


def _points(e):
    return int(np.prod(e[1])) * e[3]


def _cfg(entry, inverse=False):
    kind, shape, axes, batch, prec, _ = entry
    cfg = (FftConfig.inverse_nd(shape) if inverse or kind == "c2r" else FftConfig.new_nd(shape))
    if axes is not None:
        cfg = cfg.with_axes(axes)
    return cfg.with_batch(batch).with_precision(PREC[prec])


def _plan(entry, inverse=False):
    kind = entry[0]
    cfg = _cfg(entry, inverse)
    return {"c2c": FftPlan.c2c, "r2c": FftPlan.r2c, "c2r": FftPlan.c2r}[kind](cfg)


@pytest.fixture(scope="module")
def pool_keys(torch_cuda):
    """tuning keys of every launch of every pool plan (defaults, entries of `extra` only)."""
    keys = []
    for i, e in enumerate(sorted(POOL, key=_points)):
        W.clear_tuning()
        if e[5]:
            W.parse_tuning(e[5])
        try:
            plan = _plan(e)
        except W.FftError as err:
            print(f"pool plan {e} not built: {err}", flush=True)
            continue
        for li in range(plan.launch_count()):
            k, _ = plan.launch_tuning(li)
            if k:
                keys.append((k, e))
        del plan
    W.clear_tuning()
    return keys


def _inputs(entry, seed):
    kind, shape, axes, batch, prec, _ = entry
    cnt = int(np.prod(shape)) * batch
    if kind == "r2c":
        return oracle.normal_f32(seed, cnt)
    if kind == "c2r":
        real = oracle.normal_f32(seed, cnt)
        return oracle.r2c_nd(real, shape, batch=batch, nthreads=8).astype(np.complex64)
    x = seeded_complex(seed, cnt)
    if prec == "f32":
        return x
    return x.astype(np.complex128)


def _run(plan, entry, x):
    import torch
    prec = entry[4]
    if prec == "df64":
        dx = to_dev(oracle.df64_split(x.view(np.float64)))
    else:
        dx = to_dev(x)
    out = empty_dev(plan.required_output_buffer_size_bytes())
    plan.execute(dx, out)
    torch.cuda.synchronize()
    kind = entry[0]
    if kind == "c2r":
        return from_dev(out, np.float32)
    if prec == "f32":
        return from_dev(out, np.complex64)
    if prec == "f64":
        return from_dev(out, np.complex128)
    return oracle.df64_join(from_dev(out, np.float32)).view(np.complex128)


def _expected(entry, x, inverse):
    kind, shape, axes, batch, prec, _ = entry
    if kind == "r2c":
        return oracle.r2c_nd(x, shape, batch=batch, nthreads=8)
    if kind == "c2r":
        return oracle.c2r_nd(x.astype(np.complex128), shape, batch=batch, nthreads=8)
    return c2c_expected(x, shape, axes=axes, batch=batch, inverse=inverse)


@pytest.mark.parametrize("name", NAMES)
def test_variant_matches_oracle(pool_keys, name):
    v = parse_name(name)
    hits = [(k, e) for k, e in pool_keys if key_matches(v, k)]
    if v["pipe"]:
        # a persistent grid has `pipe` workgroups per CU (256 CUs): keep plans
        # whose pass has more tiles than that, so that workgroups loop
        pipe = int(re.search(r":pipe(\d)", name).group(1))
        hits = [(k, e) for k, e in hits if (1 << (int(k.split("|")[6][len("lines2^"):]) - 1)) // v["lines"] >= 256 * pipe]
    assert hits, f"no pool plan has a pass of the shape of {name}"
    key, entry = hits[0]
    try:
        W.clear_tuning()
        W.parse_tuning(entry[5] + f"{key} = {name}\n")
        directions = (False, True) if entry[0] == "c2c" else (False,)
        x = _inputs(entry, 500 + NAMES.index(name))
        for inverse in directions:
            plan = _plan(entry, inverse)
            used = [plan.launch_tuning(li) for li in range(plan.launch_count())]
            assert (key, name) in used, (name, key, used)
            y = _run(plan, entry, x)
            want = _expected(entry, x, inverse)
            e = oracle.rel_l2(y, want)
            print(f"{name} on {entry[0]} {entry[1]} axes={entry[2]} b={entry[3]} {entry[4]} inverse={inverse}: "
                  f"rel_l2={e:.3e} [{plan.describe()}]", flush=True)
            assert e <= TOL[entry[4]], (name, inverse, e)
    finally:
        W.clear_tuning()
