"""Host-side tuning table of the autotuner (no GPU work): text format,
save/load round trip, atomic rejection of malformed input."""
import pytest

import wgpufft_amd as W

K1 = "f32|n4096|lf00|m0/0|p0|pair0|lines2^11"
V1 = "reg:f32:n4096:t512:l1:lf00:m0/0:p0:v1:8x8x8x8"
K2 = "f32|n1000|lf00|m0/0|p0|pair0|lines2^4"


@pytest.fixture(autouse=True)
def _clean():
    W.clear_tuning()
    yield
    W.clear_tuning()


def test_parse_save_load_round_trip(tmp_path):
    n = W.parse_tuning(f"# comment\n\n{K1} = {V1}\n  {K2} =   lds:lpw=4  \n")
    assert n == 2 and W.tuning_size() == 2
    p = tmp_path / "tuning.txt"
    W.save_tuning(str(p))
    text = p.read_text()
    assert f"{K1} = {V1}" in text and f"{K2} = lds:lpw=4" in text
    W.clear_tuning()
    assert W.tuning_size() == 0
    W.load_tuning(str(p))
    assert W.tuning_size() == 2
    # re-parsing the same keys replaces, it does not duplicate
    W.parse_tuning(f"{K2} = lds:lpw=8\n")
    assert W.tuning_size() == 2


@pytest.mark.parametrize("bad", [
    f"{K1} {V1}",                                  # no '='
    f"{K1} = reg:f32:n4096:t999:l1:lf00:m0/0:p0:v1:16x16x16",  # unregistered variant
    f"{K1} = lds:lpw=0",
    f"{K1} = lds:lpw=x",
    f"{K1} = lds:lpw=99999",
    f"c64|n8 = {V1}",                              # unknown precision prefix
    "split|f32|n65536 = 4",                        # splits are two or three passes
    "fold|f32|n4096x2048 = 3",                     # the fold factor is 2 or 4 (0: off)
    "fold|f32|c2048x2048 = 4",                     # the C2R unfold is on (2) or off (0)
    "fold|f32|r2048x2048 = 4",                     # the R2C fold has radix-2 kernels only
    "fold|f32|h1048576 = 1",                       # the half-length real axis is on (2) or off (0)
])
def test_malformed_rejected_atomically(bad):
    W.parse_tuning(f"{K2} = lds:lpw=2\n")
    with pytest.raises(W.FftError) as e:
        W.parse_tuning(f"{K1} = {V1}\n{bad}\n")
    assert e.value.kind == "InvalidTuning"
    assert W.tuning_size() == 1  # nothing of the failing text was added


def test_load_missing_file(tmp_path):
    with pytest.raises(W.FftError) as e:
        W.load_tuning(str(tmp_path / "absent.txt"))
    assert e.value.kind == "InvalidTuning"


def test_split_entries():
    assert W.parse_tuning("split|f64|n16777216 = 3\nsplit|f32|n65536 = 2\n") == 2
    assert W.tuning_size() == 2


def test_fold_entries():
    text = ("fold|f32|n4096x2048 = 0\n"
            "fold|f32|n4096x256 = 4\n"
            "f32|n4096|lf00|m0/0|p0|pair0|lines2^11|fold2 = reg:f32:n4096:t512:l2:lf00:m0/0:p0:v2:8x8x8x8:fold\n"
            "fold|f32|c2048x2048 = 2\nfold|f32|r2048x2048 = 0\nfold|f32|h1048576 = 0\n")
    assert W.parse_tuning(text) == 6
    assert W.tuning_size() == 6
