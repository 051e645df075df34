"""The configurations of the reference's own GPU tests (tests/gpu_fused_pow2.rs,
gpu_fused_prime.rs, gpu_four_step.rs, gpu_small_volume.rs,
gpu_dispatch_split.rs), run through this engine's plans and checked against
the FP64 oracle at the reference's acceptance bound for f32 results (relative
L2 <= 1e-6).  The reference compares its own kernels against each other and a
CPU plan; the values are what those comparisons pin."""
import numpy as np
import pytest

import oracle
from tests.helpers import c2c_expected, empty_dev, from_dev, seeded_complex, to_dev
from wgpufft_amd import FftConfig, FftPlan, FftTuning, Normalization as Nm

pytestmark = pytest.mark.gpu

N, F, I, O = Nm.NONE, Nm.Forward, Nm.Inverse, Nm.Orthogonal

# (source, shape, batch, inverse, normalization, axes or None, tuning or None)
CASES = [
    # gpu_fused_pow2.rs
    ("fused_pow2", [2187], 1, False, N, None, None), ("fused_pow2", [2187], 1, True, I, None, None),
    ("fused_pow2", [256], 3, True, I, None, None), ("fused_pow2", [315], 3, True, I, None, None),
    ("fused_pow2", [4096], 1, False, N, None, None), ("fused_pow2", [1001], 1, False, N, None, None),
    ("fused_pow2", [13], 1, False, N, None, None), ("fused_pow2", [2048], 1, False, N, None, None),
    ("fused_pow2", [256, 12], 1, True, I, None, None), ("fused_pow2", [3, 256, 5], 1, False, I, None, None),
    ("fused_pow2", [3, 256, 5], 1, True, I, None, None), ("fused_pow2", [8, 16], 1, True, I, None, None),
    ("fused_pow2", [9, 25], 1, False, I, None, None),
    # gpu_fused_prime.rs
    ("fused_prime", [1024, 419], 1, True, I, [1], None), ("fused_prime", [1027, 1087], 1, True, I, [1], None),
    ("fused_prime", [2, 517, 3], 1, False, N, None, None), ("fused_prime", [6, 43, 5], 1, False, O, None, None),
    ("fused_prime", [6, 43, 5], 1, True, I, None, None), ("fused_prime", [1024, 986], 1, False, I, None, None),
    ("fused_prime", [2048, 179], 1, False, I, None, None), ("fused_prime", [64, 101], 1, False, I, None, None),
    ("fused_prime", [64, 613], 1, False, I, None, None), ("fused_prime", [769, 811], 1, False, I, None, None),
    ("fused_prime", [101], 1024, False, N, None, None), ("fused_prime", [101], 1024, True, I, None, None),
    ("fused_prime", [107], 1024, False, N, None, None), ("fused_prime", [1381], 2, True, I, None, None),
    ("fused_prime", [2062], 1, False, N, None, None), ("fused_prime", [2999], 1, True, I, None, None),
    ("fused_prime", [4093], 1, False, N, None, None), ("fused_prime", [4241], 2, False, N, None, None),
    ("fused_prime", [5003], 1, True, I, None, None), ("fused_prime", [613], 3, False, N, None, None),
    ("fused_prime", [613], 3, True, I, None, None),
    # gpu_four_step.rs
    ("four_step", [15, 14], 3, True, I, None, None), ("four_step", [15, 14], 1, False, N, None, None),
    ("four_step", [2, 101, 3], 1, True, I, None, None), ("four_step", [2, 101, 3], 1, False, N, None, None),
    ("four_step", [3, 5, 7, 11], 1, False, I, None, None), ("four_step", [3, 5, 7, 11], 1, True, I, None, None),
    ("four_step", [5, 17, 7], 3, True, I, None, None), ("four_step", [5, 7, 34], 2, True, I, None, None),
    ("four_step", [5, 7, 9], 3, True, I, None, None),
    # gpu_small_volume.rs
    ("small_volume", [17, 17], 5, True, I, None, None), ("small_volume", [64, 64, 32], 2, True, I, None, None),
    ("small_volume", [64, 64], 3, True, I, None, None), ("small_volume", [16, 16, 16], 1, False, N, None, None),
    ("small_volume", [16, 16, 256], 1, False, O, None, None), ("small_volume", [2, 101], 1, False, I, None, None),
    ("small_volume", [24, 20, 10], 1, False, N, None, None), ("small_volume", [31, 31], 1, False, N, None, None),
    ("small_volume", [31, 31], 1, False, I, None, FftTuning().with_direct_max_prime(0)),
    ("small_volume", [32, 16, 16], 1, False, N, None, None), ("small_volume", [47, 47], 1, False, I, None, None),
    ("small_volume", [51, 8, 3], 1, False, F, None, None), ("small_volume", [64, 64], 1, False, I, [1], None),
    ("small_volume", [8, 8, 8, 16], 1, False, N, None, None), ("small_volume", [8, 9, 10], 1, False, O, None, None),
    # gpu_dispatch_split.rs: a Bluestein line count above 65535 workgroups
    ("dispatch_split", [34], 62601, False, I, None, None), ("dispatch_split", [34], 62601, True, I, None, None),
]


@pytest.mark.parametrize("case", range(len(CASES)), ids=[f"{c[0]}-{c[1]}x{c[2]}{'-inv' if c[3] else ''}"
                                                          for c in CASES])
def test_reference_configuration(torch_cuda, case):
    import torch
    src, shape, batch, inverse, norm, axes, tuning = CASES[case]
    cfg = (FftConfig.inverse_nd(shape) if inverse else FftConfig.new_nd(shape)).with_batch(batch)
    cfg = cfg.with_normalization(norm)
    if axes is not None:
        cfg = cfg.with_axes(axes)
    if tuning is not None:
        cfg = cfg.with_tuning(tuning)
    cnt = int(np.prod(shape)) * batch
    x = seeded_complex(1000 + case, cnt)
    plan = FftPlan.c2c(cfg)
    d = to_dev(x)
    out = empty_dev(plan.required_output_buffer_size_bytes())
    plan.execute(d, out)
    torch.cuda.synchronize()
    y = from_dev(out, np.complex64)
    ref = c2c_expected(x, shape, axes=axes, batch=batch, inverse=inverse, norm=int(norm))
    e = oracle.rel_l2(y, ref)
    print(f"{src} {shape} b={batch} inverse={inverse} norm={norm.name} axes={axes} rel_l2={e:.3e} "
          f"[{plan.describe()}]", flush=True)
    assert e <= 1e-6
