"""Autotune the configurations of the benchmark workloads and save the table.

    python -m wgpufft_amd.tune [--out wgpufft_amd/tuning/mi355x.txt] [--reps 40] [--only TEXT]

Runs mifft_autotune_sequence on each benchmark workload below, as the
benchmark runs it (every candidate kernel of each line-pass shape, the whole
plan sequence timed streaming from HBM, three interleaved trials), and writes
the process-wide tuning table in its text form.  Plans consult a table
only after it is loaded (`wgpufft_amd.load_tuning`); bench.py loads the
committed table by default.
"""
from __future__ import annotations

import argparse
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_TABLE = os.path.join(HERE, "tuning", "mi355x.txt")


def workloads():
    """The benchmark workloads, each as the plan sequence the benchmark runs."""
    from . import FftConfig, FftPrecision as P, FftTransformKind as K

    def pair(shape, prec=P.F32):
        return [(FftConfig.new_nd(shape).with_precision(prec), K.C2c),
                (FftConfig.inverse_nd(shape).with_precision(prec), K.C2c)]
    return [
        ("2D C2C f32 [4096,2048] forward+inverse", pair([4096, 2048])),
        ("1D C2C f32 1024 x 4096 forward", [(FftConfig.new(1024).with_batch(4096), K.C2c)]),
        ("2D R2C then C2R f32 [2048,2048]", [(FftConfig.new_nd([2048, 2048]), K.R2c),
                                             (FftConfig.inverse_nd([2048, 2048]), K.C2r)]),
        ("3D C2C f64 [256,256,256] forward+inverse", pair([256, 256, 256], P.F64)),
        ("3D C2C df64 [256,256,256] forward+inverse", pair([256, 256, 256], P.Df64)),
        ("1D C2C f64 [2^24] forward+inverse", pair([1 << 24], P.F64)),
        ("1D C2C df64 [2^24] forward+inverse", pair([1 << 24], P.Df64)),
        # the reference's other published pair configurations (CHANGELOG.md)
        ("ref: 2D C2C f32 [2048,2048] forward+inverse", pair([2048, 2048])),
        ("ref: 2D C2C f32 [4096,4096] forward+inverse", pair([4096, 4096])),
        ("ref: 2D C2C f32 [1024,1024] forward+inverse", pair([1024, 1024])),
        ("ref: 3D C2C f32 [128,128,64] forward+inverse", pair([128, 128, 64])),
        ("ref: 3D C2C f32 [256,256,128] forward+inverse", pair([256, 256, 128])),
        ("ref: 1D C2C f32 [2^19] forward+inverse", pair([1 << 19])),
        ("ref: 1D C2C f32 [2^20] forward+inverse", pair([1 << 20])),
        ("ref: 1D C2C f32 [2^21] forward+inverse", pair([1 << 21])),
    ]


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_TABLE)
    ap.add_argument("--reps", type=int, default=40)
    ap.add_argument("--fresh", action="store_true",
                    help="start from an empty table instead of the one at --out (whose entries the "
                         "tuning does not cover, such as split choices, are otherwise kept)")
    ap.add_argument("--only", default=None,
                    help="tune only the workloads whose name contains this text (the table's other "
                         "entries are kept)")
    args = ap.parse_args(argv)
    import torch  # device runtime set-up, as in every other entry point

    torch.cuda.init()
    from . import autotune_sequence, clear_tuning, load_tuning, save_tuning, tuning_size
    clear_tuning()
    if not args.fresh and os.path.exists(args.out):
        load_tuning(args.out)
    for name, jobs in workloads():
        if args.only is not None and args.only not in name:
            continue
        n, report = autotune_sequence(jobs, reps=args.reps)
        print(f"== {name}: {n} shape(s) tuned", flush=True)
        print(report, end="", flush=True)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    save_tuning(args.out)
    print(f"saved {tuning_size()} entries to {args.out}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
