"""The plans `bench.py --workload reference` times — the configurations the
reference publishes pair timings for (its CHANGELOG.md and README.md; listed
in bench.py's REFERENCE_CONFIGS) — built with the committed MI355X tuning
table as that bench builds them, checked against the FP64 oracle: forward
and inverse of each shape (the 1D "over 1 GiB" line sets at a few MiB of
lines, the same per-line route), relative L2 <= 1e-6
(tests/gpu_accuracy.rs:94-104)."""
import importlib.util
import os

import numpy as np
import pytest

import oracle
from tests.helpers import c2c_expected, empty_dev, from_dev, seeded_complex, to_dev
import wgpufft_amd as W
from wgpufft_amd import FftConfig, FftPlan

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TABLE = os.path.join(ROOT, "wgpufft_amd", "tuning", "mi355x.txt")


def _configs():
    spec = importlib.util.spec_from_file_location("bench_configs", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.REFERENCE_CONFIGS


CONFIGS = _configs()


@pytest.fixture(autouse=True)
def _table(torch_cuda):
    W.clear_tuning()
    W.load_tuning(TABLE)
    yield
    W.clear_tuning()


@pytest.mark.parametrize("name,shape,batch", [(c[0], c[1], c[2]) for c in CONFIGS], ids=[c[0] for c in CONFIGS])
def test_published_config_matches_oracle(name, shape, batch):
    import torch
    per = int(np.prod(shape))
    b = batch if batch == 1 else max(1, min(batch, (1 << 22) // per))
    full = FftPlan.c2c(FftConfig.new_nd(shape).with_batch(batch))
    x = seeded_complex(5000 + per % 997, per * b)
    for inverse in (False, True):
        cfg = (FftConfig.inverse_nd(shape) if inverse else FftConfig.new_nd(shape)).with_batch(b)
        plan = FftPlan.c2c(cfg)
        if not inverse:
            # the same route per pass (kernel family and length) as the
            # full-size plan bench.py times
            route = lambda d: [op.split()[:2] for op in d.split(";")]
            assert route(plan.describe()) == route(full.describe()), (plan.describe(), full.describe())
        d = to_dev(x)
        out = empty_dev(x.nbytes)
        plan.execute(d, out)
        torch.cuda.synchronize()
        y = from_dev(out, np.complex64)
        want = c2c_expected(x, shape, batch=b, inverse=inverse, nthreads=16)
        e = oracle.rel_l2(y, want)
        print(f"{name} b={b} inverse={inverse} rel_l2={e:.3e} [{plan.describe()}]", flush=True)
        assert e <= 1e-6, (name, inverse, e)
