"""Benchmark of the FFT hot path on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload c2c2d|c1|r2c|c4a_f64|c4a_df64|c4b_f64|batched|serve]
                    [--gpus 1]

Default workload: 2D [4096, 2048] f32 C2C forward then inverse, in place (one
"step" = one forward+inverse pair).  Algorithmic traffic per pair is
4 * S per FFT (one HBM read + write of the 64 MiB array per axis pass, two
passes) = 536,870,912 bytes; GFLOP/s counts 5*N*log2(N) per FFT.

The inputs rotate over 8 independent 64 MiB buffers (512 MiB > the 256 MB
MALL), so every pass streams from HBM; `value` is that HBM-resident rate.  The
single-buffer (cache-resident, as the reference's README figure) rate is
reported beside it.  The roofline entry times every launch of one step live,
inside the step sequence, with HIP events on the stream it is launched on
(events that skip the system-scope cache fence; default events' figures
beside them).  The CPU baseline times the host plans (mifft_cpu_*, the
reference's CpuFftPlan restated in C++) on the same workload on the host
cores.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="c2c2d",
                    choices=["c2c2d", "c1", "r2c", "c4a_f64", "c4a_df64", "c4b_f64", "batched", "serve", "reference"])
    ap.add_argument("--requests", type=int, default=6000, help="serve workload: trace length")
    ap.add_argument("--rate", type=float, default=20000.0, help="serve workload: Poisson arrival rate (req/s)")
    ap.add_argument("--tick", type=int, default=256, help="serve workload: most requests admitted per tick")
    ap.add_argument("--no-serve", action="store_true", help="default workload: skip the serving summary")
    ap.add_argument("--transforms", type=int, default=2000, help="batched workload: transform count")
    ap.add_argument("--nbuf", type=int, default=8)
    ap.add_argument("--configs", default=None,
                    help="reference workload: comma-separated configuration names to run (default all)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--tuning", default="table", choices=["table", "auto", "none"],
                    help="kernel choices: the committed MI355X tuning table (wgpufft_amd/tuning/mi355x.txt, "
                         "made by python -m wgpufft_amd.tune), a fresh autotune of this workload, or the "
                         "planner's defaults")
    ap.add_argument("--calibrate", action="store_true",
                    help="counter driver: 5 plain 64 MiB device copies (torch), then 2 x nbuf steps "
                         "of the timed sequence; prints the launches of one step")
    return ap.parse_args()


def workload_spec(name):
    if name == "c2c2d":
        shape, batch = [4096, 2048], 1
        n_elems = 4096 * 2048
        return dict(shape=shape, batch=batch, kind="c2c_pair", elems=n_elems,
                    bytes_per_step=4 * 8 * n_elems * 2,  # 4*S per FFT, two FFTs
                    flops_per_step=2 * 5 * n_elems * 23,
                    label="2D C2C f32 [4096,2048] forward+inverse in place")
    if name == "c1":
        n, b = 1024, 4096
        return dict(shape=[n], batch=b, kind="c2c_fwd", elems=n * b, bytes_per_step=2 * 8 * n * b,
                    flops_per_step=5 * n * 10 * b,
                    label="1D C2C f32 N=1024 batch=4096 forward")
    if name in ("c4a_f64", "c4a_df64", "c4b_f64"):
        # 3D 256^3 (three axis passes) or 1D 2^24 (counted as two read+write
        # sweeps of the array per FFT, the traffic of a two-pass four-step,
        # whichever split the plan uses); one complex element is 16 bytes in
        # both f64 and df64 [re_hi,re_lo,im_hi,im_lo]
        shape = [256, 256, 256] if name.startswith("c4a") else [1 << 24]
        n_elems = 1 << 24
        passes = 3 if name.startswith("c4a") else 2
        prec = "df64" if name.endswith("df64") else "f64"
        return dict(shape=shape, batch=1, kind="c2c_pair", elems=n_elems, precision=prec, elem_bytes=16,
                    bytes_per_step=2 * passes * 16 * n_elems * 2,  # 2*P*S per FFT, two FFTs
                    flops_per_step=2 * 5 * n_elems * 24,
                    label=f"{'3D C2C [256,256,256]' if passes == 3 else '1D C2C [2^24]'} {prec} "
                          "forward+inverse in place")
    if name == "r2c":
        n = 2048
        s_r, s_p = 4 * n * n, 8 * (n // 2 + 1) * n
        return dict(shape=[n, n], batch=1, kind="r2c_c2r", elems=n * n,
                    bytes_per_step=2 * (s_r + 3 * s_p),
                    flops_per_step=2 * int(2.5 * n * n * 22),
                    label="2D R2C then C2R f32 [2048,2048]")
    raise ValueError(name)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): the library has no "
              "multi-GPU path; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from wgpufft_amd import (FftConfig, FftPlan, FftPrecision, FftTransformKind, autotune_sequence, clear_tuning,
                             fill_normal_f32, fill_uniform_f32, load_tuning, tuning_size)
    from wgpufft_amd.tune import DEFAULT_TABLE

    if args.workload == "batched":
        print(json.dumps(run_batched(args)), flush=True)
        return
    if args.workload == "serve":
        print(json.dumps(run_serve(args)), flush=True)
        return
    if args.workload == "reference":
        print(json.dumps(run_reference_configs(args)), flush=True)
        return
    spec = workload_spec(args.workload)
    torch.cuda.init()
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()

    shape, batch = spec["shape"], spec["batch"]
    plans = []
    prec = {"f32": FftPrecision.F32, "f64": FftPrecision.F64,
            "df64": FftPrecision.Df64}[spec.get("precision", "f32")]
    # kernel choices: committed tuning table, fresh autotune, or defaults
    clear_tuning()
    mode = args.tuning
    if mode == "table" and not os.path.exists(DEFAULT_TABLE):
        mode = "auto"
    tuned = {"mode": mode}
    if mode == "table":
        load_tuning(DEFAULT_TABLE)
        tuned.update(path=os.path.relpath(DEFAULT_TABLE, ROOT), entries=tuning_size())
    if mode == "auto":
        t0 = time.perf_counter()
        if spec["kind"] == "r2c_c2r":
            jobs = [(FftConfig.new_nd(shape), FftTransformKind.R2c), (FftConfig.inverse_nd(shape), FftTransformKind.C2r)]
        elif spec["kind"] == "c2c_pair":
            jobs = [(FftConfig.new_nd(shape).with_batch(batch).with_precision(prec), FftTransformKind.C2c),
                    (FftConfig.inverse_nd(shape).with_batch(batch).with_precision(prec), FftTransformKind.C2c)]
        else:
            jobs = [(FftConfig.new_nd(shape).with_batch(batch).with_precision(prec), FftTransformKind.C2c)]
        report = autotune_sequence(jobs, reps=20, stream=stream)[1]
        tuned.update(seconds=round(time.perf_counter() - t0, 3),
                     choices=[ln.split(" -> ")[0].split(":")[0] + " -> " + ln.split(" -> ")[1]
                              for ln in report.strip().splitlines() if " -> " in ln])
    if spec["kind"] == "c2c_pair":
        fwd = FftPlan.c2c(FftConfig.new_nd(shape).with_batch(batch).with_precision(prec))
        inv = FftPlan.c2c(FftConfig.inverse_nd(shape).with_batch(batch).with_precision(prec))
        plans = [fwd, inv]
        nbytes = fwd.required_buffer_size_bytes()
    elif spec["kind"] == "c2c_fwd":
        fwd = FftPlan.c2c(FftConfig.new_nd(shape).with_batch(batch))
        plans = [fwd]
        nbytes = fwd.required_buffer_size_bytes()
    else:
        r2c = FftPlan.r2c(FftConfig.new_nd(shape))
        c2r = FftPlan.c2r(FftConfig.inverse_nd(shape))
        plans = [r2c, c2r]
        nbytes = r2c.required_input_buffer_size_bytes()

    nbuf = max(1, args.nbuf)
    bufs = [torch.empty(nbytes, dtype=torch.uint8, device=dev) for _ in range(nbuf)]
    for i, b in enumerate(bufs):
        if spec["kind"] == "r2c_c2r":
            fill_normal_f32(b, 3 + i)
        elif spec.get("precision", "f32") == "f32":
            fill_uniform_f32(b, 2 + i)
        else:
            # uniform[-1,1) f32 draws, widened to f64 or placed in the df64 hi words (lo = 0)
            u = torch.empty(2 * spec["elems"] * 4, dtype=torch.uint8, device=dev)
            fill_uniform_f32(u, 5 + i)
            uf = u.view(torch.float32)
            if spec["precision"] == "f64":
                b.view(torch.float64).copy_(uf)
            else:
                hl = b.view(torch.float32).view(-1, 2)
                hl[:, 0].copy_(uf)
                hl[:, 1].zero_()
            del u, uf
    aux = None
    if spec["kind"] == "r2c_c2r":
        aux = [torch.empty(plans[0].required_output_buffer_size_bytes(), dtype=torch.uint8, device=dev)
               for _ in range(nbuf)]
    torch.cuda.synchronize()

    def step(i, single=False):
        k = 0 if single else i % nbuf
        b = bufs[k]
        if spec["kind"] == "c2c_pair":
            plans[0].execute_in_place(b, stream)
            plans[1].execute_in_place(b, stream)
        elif spec["kind"] == "c2c_fwd":
            plans[0].execute_in_place(b, stream)
        else:
            plans[0].execute(b, aux[k], stream)
            plans[1].execute(aux[k], b, stream)

    def timed(steps, single):
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for i in range(steps):
            step(i, single)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps  # ms per step

    if args.calibrate:
        # Counter calibration: a plain copy reads and writes exactly 64 MiB.
        src = torch.empty(64 << 20, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        fill_uniform_f32(src, 99)
        for _ in range(5):
            dst.copy_(src)
        # then 2 * nbuf steps of the timed sequence itself (rotating buffers)
        for i in range(2 * nbuf):
            step(i)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate": True, "copies": 5, "steps": 2 * nbuf,
                          "launches_per_step": [p.launch_describe(i) for p in plans
                                                for i in range(p.launch_count())]}))
        return

    # Correctness of what is timed (outside the timed region): the forward
    # transform of buffer 0 against torch.fft in complex128 on the device (an
    # FP64 reference independent of this library), and for pairs the round trip
    # against the input.
    fwd_err, rt_err = forward_check(spec, plans, bufs, aux, stream)
    for i in range(args.warmup):
        step(i)
    torch.cuda.synchronize()
    # median of 5 timed runs of K steps (rotating buffers), then the
    # cache-resident single-buffer rate the same way
    runs = sorted(timed(args.steps, single=False) for _ in range(5))
    ms = runs[2]
    ms_single = sorted(timed(args.steps, single=True) for _ in range(5))[2]
    value = spec["bytes_per_step"] / (ms * 1e-3) / 1e9
    value_single = spec["bytes_per_step"] / (ms_single * 1e-3) / 1e9
    gflops = spec["flops_per_step"] / (ms * 1e-3) / 1e9

    # per-launch durations inside the timed sequence: every launch of one step
    # (all plans, in order, rotating buffers) bracketed by HIP events on the
    # launch stream, averaged over `reps` steps
    kernels = in_sequence_kernel_times(spec, plans, bufs, aux, stream, nbuf, max(20, min(200, args.steps)))
    dom = max((k for k in kernels if k.get("alg_bytes")), key=lambda k: k["us"])
    traffic, traffic_src = profiled_traffic(dom["desc"], spec.get("precision", "f32"))
    pair_bytes, pair_src = profiled_step_traffic(args.workload)
    roofline = {"bound": "hbm", "achieved": round(dom["gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(dom["gbs"] / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": traffic_src,
                "kernel": dom["desc"], "kernel_us": round(dom["us"], 2),
                "alg_bytes_per_launch": dom["alg_bytes"],
                "kernel_us_default_events": dom["us_default_events"],
                "timing": f"HIP events ({dom['timing']}) around each launch inside the timed step sequence",
                "step_counter_bytes": pair_bytes, "step_counter_source": pair_src}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(spec)

    # second half of the metric: the serving workload's latency and throughput
    serve = None
    if args.workload == "c2c2d" and not args.no_serve:
        del bufs, aux
        torch.cuda.empty_cache()
        sv = run_serve(args)
        cbm, sqm = sv["modes"]["continuous_batching"], sv["modes"]["sequential_per_plan"]
        serve = {k: sv[k] for k in ("p50_ms", "p99_ms", "throughput_rps", "saturation_throughput_rps",
                                    "saturation_speedup_vs_sequential", "arrival_rate_rps", "requests")}
        serve["batched_fraction"] = cbm["poisson"]["batched_fraction"]
        serve["sequential_p50_ms"] = round(sqm["poisson"]["p50_ms"], 4)
        serve["sequential_p99_ms"] = round(sqm["poisson"]["p99_ms"], 4)
        serve["sequential_saturation_throughput_rps"] = round(sqm["saturation_throughput_rps"], 1)

    metric = "HBM GB/s (% of 8 TB/s) 4096x2048 f32 C2C; mixed-size serve p50/p99 latency"
    if args.workload != "c2c2d":
        metric = "HBM GB/s (% of 8 TB/s) " + spec["label"]
    line = {
        "metric": metric,
        "value": round(value, 1),
        "unit": "GB/s",
        "pct_of_hbm_peak": round(100 * value / HBM_PEAK_GBS, 2),
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "ms_per_step_runs": [round(r, 5) for r in runs],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": spec.get("precision", "f32"),
        "data": "synthetic (SplitMix64 uniform[-1,1) generated on device)",
        "config": {"workload": spec["label"], "shape": shape, "batch": batch,
                   "buffers": nbuf, "buffer_mib": nbytes / 2**20,
                   "alg_bytes_per_step": spec["bytes_per_step"]},
        "value_single_buffer_cache_resident": round(value_single, 1),
        "ms_per_step_single_buffer": round(ms_single, 5),
        "gflops": round(gflops, 1),
        "forward_rel_l2_vs_fp64": fwd_err,
        "round_trip_rel_l2": rt_err,
        "kernels": [{k: (round(v, 2) if isinstance(v, float) else v) for k, v in kk.items()} for kk in kernels],
        "roofline": roofline,
        "cpu_baseline": cpu,
        "autotune": tuned,
        "serve": serve,
    }
    print(json.dumps(line), flush=True)


BATCH_LENGTHS = [120, 343, 1001, 127, 1009, 4096, 65536]


# The reference's published FFT+iFFT pair timings (unstated GPU, Vulkan; the
# bench records every pair of a run into one submission on one buffer, in
# place: benches/fft_bench.rs:880-990): (name, shape, batch, published ms,
# source).  "over 1 GiB" batches the lines of a 1 GiB f32 complex buffer.
GIB_LINES = lambda n: (1 << 30) // (8 * n)
REFERENCE_CONFIGS = [
    ("4096x2048 in place", [4096, 2048], 1, 0.165, "README.md:104-106, CHANGELOG.md:79-81"),
    ("2048x2048", [2048, 2048], 1, 0.114, "CHANGELOG.md:71-73"),
    ("4096x4096", [4096, 4096], 1, 0.75, "CHANGELOG.md:74"),
    ("1024x1024", [1024, 1024], 1, 0.029, "CHANGELOG.md:85-86"),
    ("128x128x64", [128, 128, 64], 1, 0.032, "CHANGELOG.md:85"),
    ("256x256x128 in place", [256, 256, 128], 1, 0.172, "CHANGELOG.md:80-81"),
    ("65536 x 64 lines", [65536], 64, 0.079, "CHANGELOG.md:154"),
    ("2^19", [1 << 19], 1, 0.017, "CHANGELOG.md:152"),
    ("2^20", [1 << 20], 1, 0.025, "CHANGELOG.md:152"),
    ("2^21", [1 << 21], 1, 0.049, "CHANGELOG.md:153"),
    ("2^24", [1 << 24], 1, 0.78, "CHANGELOG.md:153"),
    ("17x17", [17, 17], 1, 0.0085, "CHANGELOG.md:54"),
    ("N=17 over 1 GiB", [17], GIB_LINES(17), 2.8, "CHANGELOG.md:55-56"),
    ("N=1009 over 1 GiB", [1009], GIB_LINES(1009), 3.8, "CHANGELOG.md:63-65"),
    ("N=2003 over 1 GiB", [2003], GIB_LINES(2003), 3.7, "CHANGELOG.md:66-67"),
    ("N=3256 over 1 GiB", [3256], GIB_LINES(3256), 7.4, "CHANGELOG.md:57-60"),
    ("N=4093 over 1 GiB", [4093], GIB_LINES(4093), 6.5, "CHANGELOG.md:57-60"),
    ("1229x1229", [1229, 1229], 1, 0.131, "CHANGELOG.md:69-70"),
    ("6841x6841", [6841, 6841], 1, 7.5, "CHANGELOG.md:60-61"),
]


def run_reference_configs(args):
    """The configurations the reference publishes pair timings for, measured
    as its bench runs them: forward then inverse in place on one buffer, K
    pairs recorded into one submission (a hipGraph here) and replayed, median
    of 5 replays; the forward output of each is checked against torch.fft in
    complex128."""
    import numpy as np
    import torch

    from wgpufft_amd import FftConfig, FftPlan, FftRecorder, clear_tuning, fill_uniform_f32, load_tuning
    from wgpufft_amd.tune import DEFAULT_TABLE

    torch.cuda.init()
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    clear_tuning()
    if args.tuning == "table" and os.path.exists(DEFAULT_TABLE):
        load_tuning(DEFAULT_TABLE)
    if args.calibrate:
        # counter calibration for profiles: a plain copy of exactly 64 MiB
        src = torch.empty(64 << 20, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        fill_uniform_f32(src, 99)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    rows = []
    only = set(args.configs.split(",")) if args.configs else None
    for name, shape, batch, ref_ms, src in REFERENCE_CONFIGS:
        if only is not None and name not in only:
            continue
        fwd = FftPlan.c2c(FftConfig.new_nd(shape).with_batch(batch))
        inv = FftPlan.c2c(FftConfig.inverse_nd(shape).with_batch(batch))
        nbytes = fwd.required_buffer_size_bytes()
        buf = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        fill_uniform_f32(buf, 11)
        # forward of (at most 64 MiB of) the buffer against torch.fft in complex128
        per = int(np.prod(shape))
        nchk = max(1, min(batch, (64 << 20) // (8 * per)))
        x = buf.view(torch.complex64)[:per * nchk].reshape([nchk] + shape[::-1]).to(torch.complex128)
        want = torch.fft.fftn(x, dim=list(range(1, len(shape) + 1)))
        y = buf.clone()
        fwd.execute_in_place(y, stream)
        torch.cuda.synchronize()
        got = y.view(torch.complex64)[:per * nchk].reshape([nchk] + shape[::-1]).to(torch.complex128)
        err = (torch.linalg.norm(got - want) / torch.linalg.norm(want)).item()
        del x, want, y, got

        # as the reference records every pair of a run into one command
        # buffer, the K pairs are recorded into one hipGraph (FftRecorder) on
        # a side stream and replayed; the stream-launched rate is kept beside
        def pairs(k):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(k):
                fwd.execute_in_place(buf, stream)
                inv.execute_in_place(buf, stream)
            e1.record(stream)
            e1.synchronize()
            return e0.elapsed_time(e1) / k
        first = pairs(2)
        k = int(max(3, min(200, 200.0 / max(first, 1e-3) / 5)))
        ms_stream = sorted(pairs(k) for _ in range(5))[2]
        side = torch.cuda.Stream()
        torch.cuda.synchronize()
        rec = FftRecorder(stream=side, capture_graph=True)
        for _ in range(k):
            fwd.record_in_place(rec, buf)
            inv.record_in_place(rec, buf)
        rec.end()

        def replay():
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(side)
            rec.launch(side)
            e1.record(side)
            e1.synchronize()
            return e0.elapsed_time(e1) / k
        replay()
        ms = sorted(replay() for _ in range(5))[2]
        del rec
        rows.append({"config": name, "shape": shape, "batch": batch, "ms_per_pair": round(ms, 5),
                     "ms_per_pair_stream_launched": round(ms_stream, 5),
                     "reference_ms": ref_ms, "speedup_vs_reference": round(ref_ms / ms, 2),
                     "forward_rel_l2_vs_fp64": err, "pairs": k, "route": fwd.describe(), "source": src})
        print(f"{name:22s} {ms:9.4f} ms/pair (stream-launched {ms_stream:8.4f})  reference {ref_ms:7.4f} ms  "
              f"x{ref_ms / ms:6.2f}  rel_l2 {err:.1e}",
              file=sys.stderr, flush=True)
        del buf, fwd, inv
        torch.cuda.empty_cache()
    return {"metric": "FFT+iFFT pair ms on the reference's published configurations", "unit": "ms",
            "higher_is_better": False, "n_gpus": 1, "dtype": "f32", "data": "synthetic (uniform[-1,1) on device)",
            "reference_hardware": "unstated GPU (Vulkan), README.md:402-403", "configs": rows}


def batched_items(count, seed=4):
    """The uneven-batch workload: lengths uniform over BATCH_LENGTHS, line
    counts uniform in [1, 256], 1009 forced through Bluestein."""
    import numpy as np
    rng = np.random.default_rng(seed)
    ns = rng.integers(0, len(BATCH_LENGTHS), count)
    bs = rng.integers(1, 257, count)
    return [(BATCH_LENGTHS[int(k)], int(b)) for k, b in zip(ns, bs)]


def run_batched(args):
    """One persistent batched launch (hipGraph) over `--transforms` transforms
    against sequential per-plan dispatch of the same transforms on one stream."""
    import math

    import torch

    from wgpufft_amd import FftBatch, FftConfig, FftPlan, FftTuning, fill_uniform_f32
    spec = batched_items(args.transforms)
    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream()  # graph capture needs a non-default stream
    bufs, items = [], []
    for k, (n, b) in enumerate(spec):
        t = torch.empty(n * b * 8, dtype=torch.uint8, device=dev)
        bufs.append(t)
        items.append(dict(n=n, batch=b, input=t, output=t, algorithm=2 if n == 1009 else 0))
    total_elems = sum(n * b for n, b in spec)

    def refill():
        for k, t in enumerate(bufs):
            fill_uniform_f32(t, 1000 + k, stream)

    refill()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    batch = FftBatch(items)
    t_build = time.perf_counter() - t0
    batch.capture(stream)
    info = batch.info()
    alg_bytes = sum(2 * 8 * n * b * (2 if n > 16384 else 1) for n, b in spec)
    flops = sum(5 * n * math.log2(n) * b for n, b in spec)

    def time_it(fn, reps):
        refill()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    reps = max(1, min(args.steps, 10))
    batch.launch(stream)  # warm-up
    ms_batch = time_it(lambda: batch.launch(stream), reps)
    # sequential per-plan dispatch: one standalone plan per transform, one after another
    plans = {}
    t0 = time.perf_counter()
    for it in items:
        key = (it["n"], it["batch"])
        if key not in plans:
            cfg = FftConfig.new(it["n"]).with_batch(it["batch"])
            if it["n"] == 1009:
                cfg = cfg.with_tuning(FftTuning().with_force_bluestein_axes([0]))
            plans[key] = FftPlan.c2c(cfg)
    t_plans = time.perf_counter() - t0

    def sequential():
        for it in items:
            plans[(it["n"], it["batch"])].execute_in_place(it["input"], stream)

    sequential()
    ms_seq = time_it(sequential, max(1, reps // 2))
    per_len = {}
    for n, b in spec:
        per_len.setdefault(n, [0, 0])
        per_len[n][0] += 1
        per_len[n][1] += b
    return {
        "metric": "batched uneven 1D f32 C2C: persistent launch vs sequential per-plan dispatch",
        "value": round(alg_bytes / (ms_batch * 1e-3) / 1e9, 1),
        "unit": "GB/s",
        "higher_is_better": True,
        "ms_per_launch": round(ms_batch, 3),
        "ms_sequential": round(ms_seq, 3),
        "speedup_vs_sequential": round(ms_seq / ms_batch, 2),
        "gflops": round(flops / (ms_batch * 1e-3) / 1e9, 1),
        "transforms": len(spec),
        "complex_elements": total_elems,
        "alg_bytes": alg_bytes,
        "tiles": info["tiles"], "phases": info["phases"], "grid": info["grid"],
        "lds_bytes": info["lds_bytes"], "workspace_bytes": info["workspace_bytes"],
        "batch_build_s": round(t_build, 2), "standalone_plans": len(plans), "plans_build_s": round(t_plans, 2),
        "per_length_transforms_lines": {str(k): v for k, v in sorted(per_len.items())},
        "dtype": "f32", "data": "synthetic (SplitMix64 uniform[-1,1) on device)", "n_gpus": 1,
    }


def run_serve(args):
    """Serving workload: a seeded request mix (1D f32/f64 C2C with smooth,
    prime and Bluestein lengths in [64, 2^20], 2D f32 C2C up to 1024^2, R2C)
    with Poisson arrivals at --rate; continuous batching against sequential
    per-plan dispatch, latency at the given rate and throughput at saturation
    (the whole trace arriving at once)."""
    import numpy as np
    import torch

    from wgpufft_amd import serving as S
    specs = S.request_mix(args.requests)
    bufs = S.allocate(specs)
    torch.cuda.synchronize()
    arrivals = S.poisson_arrivals(len(specs), args.rate)
    srv = S.FftServer(budget_bytes=64 << 30, streams=4, max_tick_requests=args.tick)
    t0 = time.perf_counter()
    srv.prepare(specs, bufs)
    t_prep = time.perf_counter() - t0
    srv.replay(specs[:200], bufs[:200], np.zeros(200), mode=0)  # warm-up
    out = {}
    for mode, name in ((0, "continuous_batching"), (1, "sequential_per_plan")):
        lat = S.latency_summary(srv.replay(specs, bufs, arrivals, mode=mode))
        # saturation: the whole trace at once, median of three replays
        runs = sorted((S.latency_summary(srv.replay(specs, bufs, np.zeros(len(specs)), mode=mode))
                       for _ in range(3)), key=lambda r: r["throughput_rps"])
        sat = runs[1]
        out[name] = {"poisson": lat, "saturation_throughput_rps": sat["throughput_rps"],
                     "saturation_makespan_s": sat["makespan_s"],
                     "saturation_runs_rps": [round(r["throughput_rps"], 1) for r in runs]}
    cb, sq = out["continuous_batching"], out["sequential_per_plan"]
    tot_bytes = sum(sum(S.request_bytes(s)) for s in specs)
    return {
        "metric": "mixed-size FFT serving: p50/p99 latency at Poisson arrivals, throughput at saturation",
        "value": round(cb["poisson"]["p50_ms"], 4), "unit": "ms (p50 latency)", "higher_is_better": False,
        "p50_ms": round(cb["poisson"]["p50_ms"], 4), "p99_ms": round(cb["poisson"]["p99_ms"], 4),
        "throughput_rps": round(cb["poisson"]["throughput_rps"], 1),
        "saturation_throughput_rps": round(cb["saturation_throughput_rps"], 1),
        "saturation_speedup_vs_sequential": round(cb["saturation_throughput_rps"] / sq["saturation_throughput_rps"], 2),
        "p99_ratio_sequential_over_batching": round(sq["poisson"]["p99_ms"] / cb["poisson"]["p99_ms"], 2),
        "arrival_rate_rps": args.rate, "requests": len(specs), "bytes_in_out": tot_bytes,
        "budget_bytes": 64 << 30, "streams": 4, "max_tick_requests": args.tick, "prepare_s": round(t_prep, 2),
        "modes": out, "dtype": "f32/f64", "data": "synthetic (SplitMix64 uniform[-1,1) on device)", "n_gpus": 1,
    }


def _c128(t, prec):
    """A device byte buffer of interleaved complex values as complex128."""
    import torch
    if prec == "f64":
        v = t.view(torch.float64)
    elif prec == "df64":
        hl = t.view(torch.float32).view(-1, 2).double()
        v = hl[:, 0] + hl[:, 1]
    else:
        v = t.view(torch.float32).double()
    return torch.view_as_complex(v.view(-1, 2))


def _rel(y, want):
    import torch
    return float(torch.linalg.norm((y - want).flatten()) / torch.linalg.norm(want.flatten()))


def forward_check(spec, plans, bufs, aux, stream):
    """Relative L2 of the forward transform of buffer 0 against torch.fft in
    complex128 on the device, and of the forward+inverse round trip against
    the input; buffer 0 is restored afterwards."""
    import torch
    prec = spec.get("precision", "f32")
    shape, batch = spec["shape"], spec["batch"]
    dims = list(reversed(shape))  # axis 0 is contiguous: the last torch dim
    if batch > 1:
        dims = [batch] + dims
    fft_dims = tuple(range(len(dims) - len(shape), len(dims)))
    torch.cuda.synchronize()
    ref0 = bufs[0].clone()
    if spec["kind"] == "r2c_c2r":
        plans[0].execute(bufs[0], aux[0], stream)
        torch.cuda.synchronize()
        x = ref0.view(torch.float32).double().view(*dims)
        want = torch.fft.rfftn(x, dim=fft_dims)  # halves the last dim = axis 0
        fwd = _rel(_c128(aux[0], "f32").view(want.shape), want)
        plans[1].execute(aux[0], bufs[0], stream)
        torch.cuda.synchronize()
        rt = _rel(bufs[0].view(torch.float32).double(), ref0.view(torch.float32).double())
    else:
        x = _c128(ref0, prec).view(*dims)
        plans[0].execute_in_place(bufs[0], stream)
        torch.cuda.synchronize()
        want = torch.fft.fftn(x, dim=fft_dims)
        fwd = _rel(_c128(bufs[0], prec).view(*dims), want)
        rt = None
        if spec["kind"] == "c2c_pair":
            plans[1].execute_in_place(bufs[0], stream)
            torch.cuda.synchronize()
            rt = _rel(_c128(bufs[0], prec), _c128(ref0, prec))
        del want
    bufs[0].copy_(ref0)
    torch.cuda.synchronize()
    del ref0
    return fwd, rt


HIP_EVENT_DISABLE_SYSTEM_FENCE = 0x20000000  # hip_runtime_api.h


class _FenceFreeEvents:
    """Timing events created with hipEventDisableSystemFence through the HIP
    runtime torch loaded (the same runtime libmifft.so runs on).  A default
    event performs a system-scope release when it is recorded: a writeback and
    invalidation of the caches that adds a few microseconds to every launch it
    brackets and makes the next launch start on cold caches.  These events take
    only the timestamp.  Any refusal by the runtime raises RuntimeError."""

    def __init__(self, count):
        import ctypes

        import torch
        self.ct = ctypes
        path = os.path.join(os.path.dirname(torch.__file__), "lib", "libamdhip64.so")
        hip = ctypes.CDLL(path if os.path.exists(path) else "libamdhip64.so.7")
        vp = ctypes.c_void_p
        for name, args in (("hipEventCreateWithFlags", [ctypes.POINTER(vp), ctypes.c_uint]),
                           ("hipEventRecord", [vp, vp]),
                           ("hipEventElapsedTime", [ctypes.POINTER(ctypes.c_float), vp, vp]),
                           ("hipEventDestroy", [vp])):
            fn = getattr(hip, name)
            fn.argtypes, fn.restype = args, ctypes.c_int
        self.hip, self.evs = hip, []
        for _ in range(count):
            ev = vp()
            rc = hip.hipEventCreateWithFlags(ctypes.byref(ev), HIP_EVENT_DISABLE_SYSTEM_FENCE)
            if rc != 0:
                self.close()
                raise RuntimeError(f"hipEventCreateWithFlags(hipEventDisableSystemFence) returned {rc}")
            self.evs.append(ev)

    def record(self, i, stream):
        rc = self.hip.hipEventRecord(self.evs[i], self.ct.c_void_p(int(stream.cuda_stream)))
        if rc != 0:
            raise RuntimeError(f"hipEventRecord returned {rc}")

    def elapsed_ms(self, i, j):
        ms = self.ct.c_float()
        rc = self.hip.hipEventElapsedTime(self.ct.byref(ms), self.evs[i], self.evs[j])
        if rc != 0:
            raise RuntimeError(f"hipEventElapsedTime returned {rc}")
        return ms.value

    def close(self):
        for ev in self.evs:
            self.hip.hipEventDestroy(ev)
        self.evs = []


class _TorchEvents:
    """Default HIP events (torch.cuda.Event) with the same interface."""

    def __init__(self, count):
        import torch
        self.evs = [torch.cuda.Event(enable_timing=True) for _ in range(count)]

    def record(self, i, stream):
        self.evs[i].record(stream)

    def elapsed_ms(self, i, j):
        return self.evs[i].elapsed_time(self.evs[j])

    def close(self):
        self.evs = []


def _sequence_durations(spec, plans, launches, bufs, aux, stream, nbuf, reps, events):
    """Per launch of one step, its average duration in µs over `reps` steps of
    the timed sequence (all plans in order, rotating buffers), each launch
    bracketed by events of the given kind on the launch stream."""
    import torch
    per = len(launches) + 1
    ev = events(reps * per)
    try:
        torch.cuda.synchronize()
        for r in range(reps):
            k = r % nbuf
            ev.record(r * per, stream)
            for j, (pi, li) in enumerate(launches):
                if spec["kind"] == "r2c_c2r":
                    src, dst = (bufs[k], aux[k]) if pi == 0 else (aux[k], bufs[k])
                else:
                    src = dst = bufs[k]
                plans[pi].execute_launch(li, src, dst, None, stream)
                ev.record(r * per + j + 1, stream)
        torch.cuda.synchronize()
        return [sum(ev.elapsed_ms(r * per + j, r * per + j + 1) for r in range(reps)) / reps * 1e3
                for j in range(len(launches))]
    finally:
        torch.cuda.synchronize()
        ev.close()


def in_sequence_kernel_times(spec, plans, bufs, aux, stream, nbuf, reps):
    """Average duration of every launch of one step as the timed region runs
    it (all plans in order, rotating buffers), each launch bracketed by HIP
    events on the launch stream; with its algorithmic bytes (one read and one
    write of what the pass moves).  `us` comes from events that skip the
    system-scope fence when the runtime accepts them (`timing` says which),
    `us_default_events` from default events."""
    launches = [(pi, li) for pi, p in enumerate(plans) for li in range(p.launch_count())]
    fenced = _sequence_durations(spec, plans, launches, bufs, aux, stream, nbuf, reps, _TorchEvents)
    try:
        times = _sequence_durations(spec, plans, launches, bufs, aux, stream, nbuf, reps, _FenceFreeEvents)
        timing = "hipEventDisableSystemFence"
    except (OSError, AttributeError, RuntimeError) as e:
        print(f"bench: fence-free events unavailable ({e}); timing launches with default events",
              file=sys.stderr, flush=True)
        times, timing = fenced, "default"
    es = spec.get("elem_bytes", 8)
    if spec["kind"] == "r2c_c2r":
        n = spec["shape"][0]
        s_r, s_p = 4 * spec["elems"], 8 * (n // 2 + 1) * (spec["elems"] // n)
        per = {(0, 0): s_r + s_p, (0, 1): 2 * s_p, (1, 0): 2 * s_p, (1, 1): s_p + s_r}
        if any(p.launch_count() != 2 for p in plans):
            per = {}
    else:
        per = {lp: 2 * es * spec["elems"] for lp in launches}
    out = []
    for j, (pi, li) in enumerate(launches):
        us = times[j]
        k = dict(plan=pi, launch=li, desc=plans[pi].launch_describe(li), us=us,
                 us_default_events=round(fenced[j], 2), timing=timing, alg_bytes=per.get((pi, li)))
        if k["alg_bytes"]:
            k["gbs"] = k["alg_bytes"] / (us * 1e-6) / 1e9
        out.append(k)
    return out


def profiled_step_traffic(workload):
    """HBM-side bytes (FETCH_SIZE + WRITE_SIZE, copy-calibrated) of one whole
    timed step, from the committed counter summary of the timed region, or
    (None, None)."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        for line in open(path):
            if line.startswith(f"step_traffic_bytes: workload={workload} "):
                return int(line.rsplit("=", 1)[1]), os.path.relpath(path, ROOT)
    return None, None


def profiled_traffic(desc, prec="f32"):
    """HBM bytes per launch of a kernel from the committed rocprofv3 counter
    summaries (FETCH_SIZE + WRITE_SIZE, copy-calibrated), matched on precision,
    length, radices, threads per line, vector width, line-first load / store
    and the fold, post-twiddle and pipelined forms; (None, None) when no
    summary covers it."""
    import glob
    import re
    m = re.search(r"N=(\d+) (\S+) .*tpl=(\d+) vec=(\d)", desc)
    if not m:
        return None, None
    n, rad, tpl, vec = m.groups()
    lf = "lf=" + ("t" if "ldLF" in desc else "f") + ("t" if "stLF" in desc else "f")
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        for line in open(path):
            if not line.startswith("traffic_bytes_per_launch:"):
                continue
            if (f" {prec} N={n} " in line and f" tpl={tpl} " in line and f"radices={rad} " in line
                    and f"vec={vec}" in line
                    and f" {lf} " in line
                    and (" post" in line) == (" post" in desc) and ("pipe" in line) == ("pipe" in desc)
                    and (" fold" in line) == (" fold" in desc)):
                return int(line.rsplit("=", 1)[1]), os.path.relpath(path, ROOT)
    return None, None


def cpu_baseline(spec):
    """The host-memory plans (CpuFftPlan restated in C++, mifft_cpu_*) on the
    host cores: the same workload in the same precision, one warm-up, then the
    median of 5 runs (benches/cpu_fft_bench.rs:64-84), bounded by a sample of
    at most about 20 s of CPU work."""
    import numpy as np

    from wgpufft_amd import (CpuFftPlan, FftConfig, FftPrecision, cpu_isa, cpu_threads, fill_uniform_f32)
    shape, batch = spec["shape"], spec["batch"]
    prec = spec.get("precision", "f32")
    p = {"f32": FftPrecision.F32, "f64": FftPrecision.F64, "df64": FftPrecision.Df64}[prec]
    rng = np.random.default_rng(2)
    if spec["kind"] == "r2c_c2r":
        r2c = CpuFftPlan.r2c(FftConfig.new_nd(shape))
        c2r = CpuFftPlan.c2r(FftConfig.inverse_nd(shape))
        x = rng.standard_normal(r2c.required_input_len()).astype(np.float32)
        pk = np.zeros(r2c.required_output_len(), np.float32)

        def one():
            r2c.execute(x, pk)
            c2r.execute(pk, x)
        what = "R2C+C2R pair"
    else:
        plans = [CpuFftPlan.c2c(FftConfig.new_nd(shape).with_batch(batch).with_precision(p))]
        if spec["kind"] == "c2c_pair":
            plans.append(CpuFftPlan.c2c(FftConfig.inverse_nd(shape).with_batch(batch).with_precision(p)))
        dt = np.float64 if prec == "f64" else np.float32
        x = rng.uniform(-1, 1, plans[0].required_input_len()).astype(dt)

        def one():
            for pl in plans:
                if prec == "f64":
                    pl.execute_in_place_f64(x)
                else:
                    pl.execute_in_place(x)
        what = "forward+inverse pair" if len(plans) == 2 else "forward FFT"
    t0 = time.perf_counter()
    one()  # warm-up
    first = time.perf_counter() - t0
    runs = max(3, min(25, int(10.0 / max(first, 1e-6))))
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        one()
        times.append(time.perf_counter() - t0)
    dt_s = sorted(times)[len(times) // 2]
    return {"value": round(spec["bytes_per_step"] / dt_s / 1e9, 3), "unit": "GB/s",
            "ms_per_step": round(dt_s * 1e3, 2), "cores": cpu_threads(), "kind": "port",
            "sample": f"one {what} of the same {shape} x{batch} {prec} workload per run, median of {runs} runs "
                      f"after one warm-up, through the host plans (mifft_cpu_*, CpuFftPlan restated in C++, "
                      f"{cpu_isa()}, {cpu_threads()} threads)"}


if __name__ == "__main__":
    main()
