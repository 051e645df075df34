"""CPU tests of the C ABI boundary: the library loads, exports every entry point
include/mifft.h declares, and validates configurations like the reference's
FftConfig::validate / FftTuning::validate_for_config (src/config.rs:178-229,
src/tuning.rs:280-336) -- no kernel runs here."""
import ctypes
import os
import re

import pytest

import wgpufft_amd as W
from wgpufft_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mifft.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mifft_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 35
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_signatures_cover_the_header():
    bound = {name for name, _, _ in _lib.SIGNATURES}
    assert set(declared_functions()) <= bound


def test_struct_layout_matches_header_expectations():
    # mifft_config: 8 shape u64, rank, 8 axes, naxes, batch, 3 enums, tuning
    assert ctypes.sizeof(_lib.Config) == 8 * 8 + 4 + 8 * 4 + 4 + 8 + 3 * 4 + 4 + ctypes.sizeof(_lib.Tuning)


def test_defaults_mirror_reference():
    """FftConfig::new_nd and FftTuning::default (config.rs:90-103, tuning.rs:55-77)."""
    lib = _lib.load()
    c = _lib.Config()
    shape = (ctypes.c_uint64 * 2)(4096, 2048)
    assert lib.mifft_config_init(ctypes.byref(c), 2, shape) == 0
    assert list(c.axes[:2]) == [0, 1] and c.naxes == 2 and c.batch == 1
    assert c.direction == 0 and c.normalization == 2 and c.precision == 0
    t = c.tuning
    assert (t.workgroup_size, t.fused_workgroup_size, t.rader_max_prime, t.direct_max_prime) == (64, 256, 8192, 127)
    assert t.fuse_long_axes == 1 and t.fuse_small_volumes == 1 and t.fused_min_convolution_length == 128


def err(cfg):
    with pytest.raises(W.FftError) as e:
        cfg.validate()
    return e.value.kind


def test_validation_errors():
    assert err(W.FftConfig.new_nd([])) == "ZeroLength"
    assert err(W.FftConfig.new_nd([4, 0])) == "ZeroLength"
    assert err(W.FftConfig.new(8).with_batch(0)) == "ZeroBatch"
    assert err(W.FftConfig.new(8).with_axes([])) == "EmptyAxes"
    assert err(W.FftConfig.new_nd([4, 4]).with_axes([2])) == "InvalidAxis"
    assert err(W.FftConfig.new_nd([4, 4]).with_axes([1, 1])) == "DuplicateAxis"
    assert err(W.FftConfig.new_nd([1, 4])) == "UnsupportedLength"  # a length-1 axis of a >1-point transform
    W.FftConfig.new_nd([1, 1]).validate()  # one point total is the direct route
    assert err(W.FftConfig.new_nd([65536, 65536])) == "LengthTooLarge"  # one transform > u32::MAX elements
    assert err(W.FftConfig.new(2 ** 33)) == "LengthTooLarge"
    W.FftConfig.new(2 ** 20).with_batch(2 ** 12 - 1).validate()
    # more batches than u32::MAX elements: the large-chunk route (independent
    # batch chunks), up to 2^59 elements
    W.FftConfig.new(2 ** 20).with_batch(2 ** 12).validate()
    W.FftConfig.new_nd([65536, 65535]).with_batch(3).validate()
    assert err(W.FftConfig.new(2 ** 30).with_batch(2 ** 30)) == "LengthTooLarge"


def test_tuning_validation_errors():
    base = W.FftConfig.new_nd([17, 16])
    assert err(base.with_tuning(W.FftTuning().with_workgroup_size(0))) == "InvalidTuning"
    assert err(base.with_tuning(W.FftTuning().with_workgroup_size(48))) == "InvalidTuning"
    assert err(base.with_tuning(W.FftTuning().with_fused_workgroup_size(3))) == "InvalidTuning"
    assert err(base.with_tuning(W.FftTuning().with_rader_max_prime(1))) == "InvalidTuning"
    t = W.FftTuning().with_force_rader_axes([0]).with_force_bluestein_axes([0])
    assert err(base.with_tuning(t)) == "InvalidTuning"  # conflicting
    assert err(base.with_tuning(W.FftTuning().with_force_rader_axes([1]))) == "InvalidTuning"  # 16 not prime
    assert err(base.with_tuning(W.FftTuning().with_force_rader_axes([0, 0]))) == "InvalidTuning"  # duplicate
    assert err(base.with_tuning(W.FftTuning().with_force_bluestein_axes([2]))) == "InvalidTuning"  # out of range
    assert err(base.with_axes([1]).with_tuning(W.FftTuning().with_force_rader_axes([0]))) == "InvalidTuning"
    base.with_tuning(W.FftTuning().with_force_rader_axes([0]).with_force_bluestein_axes([1])).validate()


def test_scale_uses_full_shape_product():
    """config.rs:262-286: scale = f(prod(shape)) whatever the axes."""
    c = W.FftConfig.new_nd([8, 4]).with_axes([1])
    assert c.scale_f64() == 1.0
    assert c.with_direction(W.FftDirection.Inverse).scale_f64() == 1.0 / 32
    assert c.with_normalization(W.Normalization.Forward).scale_f64() == 1.0 / 32
    assert c.with_normalization(W.Normalization.Orthogonal).scale_f64() == pytest.approx(32 ** -0.5, rel=1e-16)
    assert c.with_normalization(W.Normalization.NONE).with_direction(1).scale_f64() == 1.0


def test_sizes():
    c = W.FftConfig.new_nd([4096, 2048]).with_batch(2)
    assert c.total_complex_len() == 2 * 4096 * 2048
    assert c.required_buffer_size_bytes() == 2 * 4096 * 2048 * 8
    assert c.with_precision(W.FftPrecision.Df64).required_buffer_size_bytes() == 2 * 4096 * 2048 * 16


def test_status_strings_and_no_throw_on_null():
    lib = _lib.load()
    for code in range(0, 25):
        assert lib.mifft_status_string(code)
    assert lib.mifft_config_validate(None) == 17
    assert lib.mifft_plan_create(None, 0, None) == 17
    assert lib.mifft_execute(None, None, None, 0, None, 0) != 0
    assert b"mifft" in lib.mifft_version()


def test_registry_names_parse_and_are_unique():
    """mifft_registry_count / mifft_registry_name read the compiled variant table (no GPU call)."""
    from tests.test_gpu_registry import parse_name
    names = W.registry_names()
    assert len(names) == len(set(names)) and len(names) > 50
    for nm in names:
        parse_name(nm)
    # the committed MI355X table names only registered variants
    table = os.path.join(ROOT, "wgpufft_amd", "tuning", "mi355x.txt")
    for line in open(table):
        line = line.split("#")[0].strip()
        if "= reg:" in line:
            assert line.split("=", 1)[1].strip() in names, line
