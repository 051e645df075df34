"""GPU parity of the direct prime DFT (src/runtime/direct_prime.rs:25-110;
route c2c.rs:245-254, rader_axis.rs:290-320): f32 primes up to
direct_max_prime whose p - 1 is not {2,3,5,7,11,13}-smooth (47, 59, 83, 103,
107), on contiguous and strided axes, both directions, against the oracle;
forced Rader and a lower direct_max_prime keep the Rader route."""
import numpy as np
import pytest

import oracle
from tests.helpers import c2c_expected, empty_dev, from_dev, seeded_complex, to_dev
from wgpufft_amd import FftConfig, FftDirection, FftPlan, FftTuning

pytestmark = pytest.mark.gpu


def run(cfg, x):
    import torch
    plan = FftPlan.c2c(cfg)
    d = to_dev(x)
    out = empty_dev(x.nbytes)
    plan.execute(d, out)
    torch.cuda.synchronize()
    return plan, from_dev(out, np.complex64)


# (2D volumes above 4096 points: smaller ones take the small-volume kernel,
# whose prime axes are direct DFTs too, tests/test_gpu_small_volume.py)
@pytest.mark.parametrize("shape,batch", [([47], 300), ([59], 7), ([83], 1000), ([103], 3), ([107], 50),
                                         ([128, 59], 2), ([47, 103], 1), ([48, 107], 4)])
@pytest.mark.parametrize("direction", [FftDirection.Forward, FftDirection.Inverse])
def test_direct_prime_matches_oracle(torch_cuda, shape, batch, direction):
    n = int(np.prod(shape))
    x = seeded_complex(700 + n + batch, n * batch)
    cfg = FftConfig.new_nd(shape).with_batch(batch).with_direction(direction)
    plan, y = run(cfg, x)
    assert "direct N=" in plan.describe(), plan.describe()
    ref = c2c_expected(x, shape, batch=batch, inverse=direction == FftDirection.Inverse)
    e = oracle.rel_l2(y, ref)
    print(f"{shape} b={batch} {direction.name} rel_l2={e:.3e} max_rel={oracle.max_rel(y, ref):.3e} "
          f"[{plan.describe()}]", flush=True)
    assert e <= 1e-6


def test_direct_prime_route_selection(torch_cuda):
    x = seeded_complex(5, 59 * 8)
    forced = FftTuning().with_force_rader_axes([0])
    lower = FftTuning().with_direct_max_prime(53)
    for t in (forced, lower):
        plan, y = run(FftConfig.new(59).with_batch(8).with_tuning(t), x)
        d = plan.describe()
        # forced: Rader; above direct_max_prime: 58 is not smooth, so the
        # zero-padded Rader convolution gives way to the register Bluestein
        # kernel (rader_axis.rs:1260-1268)
        assert "direct" not in d and ("rader" in d if t is forced else "bluestein-reg" in d), d
        assert oracle.rel_l2(y, c2c_expected(x, [59], batch=8)) <= 1e-6
    # cyclic Rader (p - 1 smooth) keeps Rader
    assert "direct" not in FftPlan.c2c(FftConfig.new(61)).describe()
