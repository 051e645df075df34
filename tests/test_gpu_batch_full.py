"""The uneven-batch workload at full size (the persistent launch bench.py
times): 2000 transforms, lengths uniform over {120, 343, 1001, 127, 1009
(forced Bluestein), 4096, 65536}, 1..256 lines each, about 21 GB on the
device, one captured persistent launch.  A seeded sample of 4 lines per
transform (first, last and two drawn) is checked against the FP64 oracle per
length class (relative L2 <= 1e-6 and max-rel <= 5e-6); 24 transforms are
compared bit for bit with their standalone plans."""
import numpy as np
import pytest

import oracle
from tests.helpers import empty_dev, from_dev
import wgpufft_amd as W
from wgpufft_amd import FftBatch, FftConfig, FftPlan, FftTuning

pytestmark = pytest.mark.gpu


def test_full_uneven_batch_against_oracle(torch_cuda):
    import torch
    from bench import batched_items
    W.clear_tuning()
    spec = batched_items(2000)
    stream = torch.cuda.Stream()
    bufs, items = [], []
    for k, (n, b) in enumerate(spec):
        t = torch.empty(n * b * 8, dtype=torch.uint8, device="cuda")
        W.fill_uniform_f32(t, 1000 + k, stream)
        bufs.append(t)
        items.append(dict(n=n, batch=b, input=t, output=t, algorithm=2 if n == 1009 else 0))
    torch.cuda.synchronize()
    rng = np.random.default_rng(44)
    picks = []
    for k, (n, b) in enumerate(spec):
        lines = sorted(set([0, b - 1] + [int(v) for v in rng.integers(0, b, 2)]))
        picks.append(lines)
    xin = [from_dev(bufs[k].view(b, n * 8)[picks[k]].contiguous(), np.complex64).reshape(len(picks[k]), n).copy()
           for k, (n, b) in enumerate(spec)]
    standalone_ids = [int(v) for v in rng.choice(len(spec), 24, replace=False)]
    copies = {k: bufs[k].clone() for k in standalone_ids}
    torch.cuda.synchronize()
    batch = FftBatch(items)
    batch.capture(stream)
    batch.launch(stream)
    torch.cuda.synchronize()
    print("batch info", batch.info(), flush=True)
    per_class = {}
    for k, (n, b) in enumerate(spec):
        y = from_dev(bufs[k].view(b, n * 8)[picks[k]].contiguous(), np.complex64).reshape(len(picks[k]), n)
        per_class.setdefault(n, []).append((xin[k], y))
    for n, pairs in sorted(per_class.items()):
        x = np.concatenate([p[0] for p in pairs]).reshape(-1)
        y = np.concatenate([p[1] for p in pairs]).reshape(-1)
        ref = oracle.c2c_nd(x.astype(np.complex128), [n], batch=x.size // n, nthreads=16)
        e, m = oracle.rel_l2(y, ref), oracle.max_rel(y, ref)
        print(f"length {n}: {len(pairs)} transforms, {x.size // n} sampled lines, rel_l2={e:.3e} max_rel={m:.3e}",
              flush=True)
        assert e <= 1e-6 and m <= 5e-6, (n, e, m)
    for k in standalone_ids:
        n, b = spec[k]
        cfg = FftConfig.new(n).with_batch(b)
        if n == 1009:
            cfg = cfg.with_tuning(FftTuning().with_force_bluestein_axes([0]))
        plan = FftPlan.c2c(cfg)
        out = empty_dev(copies[k].numel())
        plan.execute(copies[k], out)
        torch.cuda.synchronize()
        assert np.array_equal(from_dev(out, np.uint8), from_dev(bufs[k], np.uint8)), (k, n, b)
    print(f"{len(standalone_ids)} transforms equal their standalone plans bit for bit", flush=True)
    del bufs, items, batch
