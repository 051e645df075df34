"""Routing and planning decision table (the reference's c2c.rs:8950-9015,
axis_policy.rs:72-177, runtime/mod.rs:29-54): for a matrix of lengths the
plan reports the axis kind, route and greedy factorization the reference's
rules give, restated here from their definitions (smooth over
{2,3,5,7,11,13} -> Mixed; prime <= rader_max_prime -> Rader; otherwise
Bluestein; greedy factors over [13,11,8,7,5,4,3,2])."""
import pytest

from wgpufft_amd import AxisKind, C2cRoute, FftConfig, FftError, FftPlan, FftTuning

pytestmark = pytest.mark.gpu


def is_prime(n):
    if n < 2:
        return False
    d = 2
    while d * d <= n:
        if n % d == 0:
            return False
        d += 1
    return True


def greedy_factors(n):
    out = []
    for r in (13, 11, 8, 7, 5, 4, 3, 2):
        while n % r == 0:
            out.append(r)
            n //= r
    return out if n == 1 else None


def expected_kind(n, rader_max=8192):
    if greedy_factors(n) is not None:
        return AxisKind.Mixed
    if is_prime(n) and n <= rader_max:
        return AxisKind.Rader
    return AxisKind.Bluestein


LENGTHS = [2, 3, 4, 5, 7, 8, 11, 13, 16, 17, 19, 26, 34, 47, 97, 120, 127, 128, 143, 169, 243, 343, 1000,
           1001, 1009, 2048, 2999, 3000, 3256, 4096, 8191, 8209, 10007, 65536, 65537]


@pytest.mark.parametrize("n", LENGTHS)
def test_axis_kind_route_and_factors(torch_cuda, n):
    plan = FftPlan.c2c(FftConfig.new(n))
    kind = expected_kind(n)
    assert plan.axis_kinds() == [kind], (n, plan.axis_kinds())
    route = {AxisKind.Mixed: C2cRoute.MixedRadix, AxisKind.Rader: C2cRoute.Rader,
             AxisKind.Bluestein: C2cRoute.Bluestein}[kind]
    assert plan.route() == route
    f = greedy_factors(n)
    if f is not None:
        assert plan.axis_factors()[0] == f


def test_tuning_changes_routes(torch_cuda):
    assert FftPlan.c2c(FftConfig.new(1009).with_tuning(FftTuning().with_rader_max_prime(1000))).axis_kinds() == \
        [AxisKind.Bluestein]
    # forced axes win over the length rule (axis_policy.rs:55-61); a forced
    # Rader axis must be a prime >= 3 (tuning.rs validate_forced_axes)
    assert FftPlan.c2c(FftConfig.new(8209).with_tuning(FftTuning().with_force_rader_axes([0]))).axis_kinds() == \
        [AxisKind.Rader]
    assert FftPlan.c2c(FftConfig.new(64).with_tuning(FftTuning().with_force_bluestein_axes([0]))).axis_kinds() == \
        [AxisKind.Bluestein]
    with pytest.raises(FftError) as e:
        FftPlan.c2c(FftConfig.new(64).with_tuning(FftTuning().with_force_rader_axes([0])))
    assert e.value.kind == "InvalidTuning"
    two = FftPlan.c2c(FftConfig.new_nd([120, 127]))
    assert two.axis_kinds() == [AxisKind.Mixed, AxisKind.Rader] and two.route() == C2cRoute.AxisSequence
    assert FftPlan.c2c(FftConfig.new_nd([64, 64])).route() == C2cRoute.MixedRadix
