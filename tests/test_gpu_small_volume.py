"""GPU parity of the small-volume route (SmallVolumePlan,
src/runtime/small_volume.rs:65-275): every axis of an f32 volume of at most
4096 points in one workgroup, prime axes up to direct_max_prime as direct
DFTs (small_volume.rs:9-10, 220-245), against the FP64 oracle; the route is
taken only where the reference takes it and switched off by
fuse_small_volumes."""
import numpy as np
import pytest

import oracle
from tests.helpers import c2c_expected, empty_dev, from_dev, seeded_complex, to_dev
from wgpufft_amd import FftConfig, FftDirection, FftPlan, FftPrecision, FftTuning, Normalization

pytestmark = pytest.mark.gpu

SHAPES = [([16, 16], 1), ([32, 64], 3), ([8, 8, 8], 2), ([16, 16, 16], 1), ([32, 128], 5), ([12, 10, 9], 7),
          ([2, 2, 2, 2, 2, 2, 2, 2], 9), ([120, 9], 4), ([4, 1024], 2), ([13, 11, 7], 3),
          # prime axes (direct DFTs inside the workgroup)
          ([17, 17], 3), ([17, 16], 2), ([31, 8, 5], 4), ([101, 11], 1), ([3, 127], 5), ([59, 47], 2)]


def run(cfg, x, in_place=False):
    import torch
    plan = FftPlan.c2c(cfg)
    d = to_dev(x)
    if in_place:
        plan.execute_in_place(d)
        out = d
    else:
        out = empty_dev(x.nbytes)
        plan.execute(d, out)
    torch.cuda.synchronize()
    return plan, from_dev(out, np.complex64)


@pytest.mark.parametrize("shape,batch", SHAPES)
@pytest.mark.parametrize("direction", [FftDirection.Forward, FftDirection.Inverse])
def test_small_volume_matches_oracle(torch_cuda, shape, batch, direction):
    n = int(np.prod(shape))
    x = seeded_complex(300 + n + batch, n * batch)
    cfg = FftConfig.new_nd(shape).with_batch(batch).with_direction(direction)
    plan, y = run(cfg, x, in_place=batch % 2 == 1)
    assert plan.describe().startswith("small-volume"), plan.describe()
    ref = c2c_expected(x, shape, batch=batch, inverse=direction == FftDirection.Inverse)
    e = oracle.rel_l2(y, ref)
    print(f"{shape} b={batch} {direction.name} rel_l2={e:.3e} [{plan.describe()}]", flush=True)
    assert e <= 1e-6


@pytest.mark.parametrize("norm", [Normalization.NONE, Normalization.Forward, Normalization.Orthogonal])
def test_small_volume_normalizations(torch_cuda, norm):
    shape = [32, 32]
    x = seeded_complex(41, 1024 * 2)
    cfg = FftConfig.new_nd(shape).with_batch(2).with_normalization(norm)
    plan, y = run(cfg, x)
    assert plan.describe().startswith("small-volume")
    ref = c2c_expected(x, shape, batch=2, norm=int(norm))
    assert oracle.rel_l2(y, ref) <= 1e-6


def test_small_volume_route_selection(torch_cuda):
    x = seeded_complex(42, 32 * 64)
    off = FftTuning().with_fuse_small_volumes(False)
    p_off, y_off = run(FftConfig.new_nd([32, 64]).with_tuning(off), x)
    p_on, y_on = run(FftConfig.new_nd([32, 64]), x)
    assert not p_off.describe().startswith("small-volume") and p_on.describe().startswith("small-volume")
    ref = c2c_expected(x, [32, 64])
    assert oracle.rel_l2(y_off, ref) <= 1e-6 and oracle.rel_l2(y_on, ref) <= 1e-6
    # not taken: more than 4096 points, a prime axis above direct_max_prime
    # or with Rader forced, a subset of axes, f64, 2D with both sides >= 64
    for cfg in (FftConfig.new_nd([128, 64]), FftConfig.new_nd([131, 16]),
                FftConfig.new_nd([17, 16]).with_tuning(FftTuning().with_force_rader_axes([0])),
                FftConfig.new_nd([17, 16]).with_tuning(FftTuning().with_direct_max_prime(13)),
                FftConfig.new_nd([16, 16]).with_axes([1]),
                FftConfig.new_nd([16, 16]).with_precision(FftPrecision.F64), FftConfig.new_nd([64, 64])):
        assert not FftPlan.c2c(cfg).describe().startswith("small-volume"), cfg
    assert FftPlan.c2c(FftConfig.new_nd([17, 16])).describe().startswith("small-volume")
