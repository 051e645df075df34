"""bench.py's host-side contract, without a GPU: the one-GPU refusal, the
algorithmic bytes of the headline workload, and the lookup of the committed
counter summaries that fills `roofline.traffic` for the launches the
4096x2048 f32 pair runs (their descriptions as the plans print them)."""
import importlib.util
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_contract", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and r.stdout == ""


def test_headline_bytes_per_pair():
    spec = _bench().workload_spec("c2c2d")
    # one read and one write of the 64 MiB array per axis pass, two passes per
    # FFT, forward and inverse
    assert spec["bytes_per_step"] == 4 * (4096 * 2048 * 8) * 2 == 536870912
    assert spec["shape"] == [4096, 2048] and spec["kind"] == "c2c_pair"


def test_traffic_lookup_for_headline_launches():
    b = _bench()
    fold = "reg N=4096 8x8x8x8 lines=2048 tpl=256 vec=2 lpw=2 fold2 M=2048"
    col = "reg N=1024 16x16x4 lines=8192 tpl=64 vec=2 lpw=8 ldLF stLF"
    t_fold, src = b.profiled_traffic(fold)
    t_col, src2 = b.profiled_traffic(col)
    assert src == src2 == os.path.join("profiles", "c2c2d_4096x2048_f32.md")
    # within a few percent of the algorithmic 134,217,728 B per pass
    for t in (t_fold, t_col):
        assert 134217728 <= t <= 1.05 * 134217728
    pair, psrc = b.profiled_step_traffic("c2c2d")
    # the pair is twice both passes (each figure rounded to whole bytes)
    assert psrc == src and abs(pair - 2 * (t_fold + t_col)) <= 4
    # a launch no summary covers gives no figure rather than a neighbour's
    assert b.profiled_traffic(fold.replace("fold2 ", "")) == (None, None)
    assert b.profiled_traffic(col.replace("tpl=64", "tpl=128")) == (None, None)
