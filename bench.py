#!/usr/bin/env python3
"""Falling-sand tick throughput on one MI355X.

Default workload: the 16384 x 16384 all-active world (268,435,456 cells,
2.1 GB per state buffer; rows alternate 50% sand / air, 10% water rows, 10%
steam rows, stone floor; reactions off; uniform 20 degrees), ticked in
place in HBM.  A step is one tick of the whole world; the timed region is
exactly --steps ticks between two stream synchronisations.

Prints ONE JSON line: Gcell-updates/s, the HBM roofline of the tick kernel
(16 algorithmic bytes per cell-update: one 8-byte cell read and one write),
the CPU baseline (the reference's sequential tick restated, timed on this
host's cores on a bounded sample), and secondary figures for the 4096^2
dirty-chunk world, the 256-world batch and the serving loop.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_UPDATE = 16


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bench_c3(fsg, steps, warmup, side=16384, seed=3):
    cells = fsg.generate(3, side, side, seed)
    cfg = fsg.default_config(reactions=0, temp_every=1)
    ctx = fsg.Context(0)
    w = fsg.World(ctx, side, side, seed, cfg)
    w.upload(cells)
    del cells
    if warmup:
        w.tick(warmup)
    w.sync()
    t0 = time.perf_counter()
    ms_dev = w.step_timed(steps)  # HIP events on the ctx stream around the launches
    w.sync()
    wall = time.perf_counter() - t0
    st = w.stats()
    w.close()
    ctx.close()
    return {"wall_s": wall, "dev_ms": ms_dev, "cells": side * side, "counts": list(st.count)[:8]}


def cpu_baseline(threads=None, side_w=4096, side_h=1024, ticks=4):
    """refseq (the reference's sequential chunk-checkerboard tick, restated)
    on a bounded sample of the all-active world."""
    import falling_sand_engine_amd as fsg
    from oracle import pyoracle
    threads = threads or min(16, os.cpu_count() or 1)
    cells = fsg.generate(3, side_w, side_h, 3)
    cfg = pyoracle.default_config(reactions=0, temp_every=1)
    t0 = time.perf_counter()
    pyoracle.refseq(cells, cfg, 3, ticks, threads)
    dt = time.perf_counter() - t0
    return {"value": side_w * side_h * ticks / dt / 1e9, "unit": "Gcell-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"refseq restatement of world::tick, all-active striped world {side_w}x{side_h}, "
                      f"{ticks} ticks, {threads} threads, {dt:.2f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--side", type=int, default=16384)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mode", default="all", choices=["all", "c3", "copy"],
                    help="all: the JSON line; c3: only the timed tick loop (for profiling); "
                         "copy: the HBM copy calibration kernel")
    ap.add_argument("--copy-bytes", type=int, default=2 << 30)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: --gpus must be 1 "
                                   "(the multi-GPU sharding is a design, not built)"}))
        sys.exit(2)
    if int(os.environ.get("RANK", "0")) != 0:
        return
    import falling_sand_engine_amd as fsg

    if args.mode == "copy":
        with fsg.Context(0) as ctx:
            ms = ctx.copy_timed(args.copy_bytes, max(1, args.steps))
        gbs = 2 * args.copy_bytes / (ms / 1e3) / 1e9
        print(json.dumps({"mode": "copy", "bytes": args.copy_bytes, "ms": ms, "GBs_read_plus_write": gbs}))
        return
    if args.mode == "c3":
        r = bench_c3(fsg, args.steps, args.warmup, args.side)
        print(json.dumps({"mode": "c3", "steps": args.steps, "dev_ms": r["dev_ms"], "wall_s": r["wall_s"]}))
        return

    log(f"[bench] C3 {args.side}^2 warmup={args.warmup} steps={args.steps}")
    r = bench_c3(fsg, args.steps, args.warmup, args.side)
    cell_updates = r["cells"] * args.steps
    value = cell_updates / r["wall_s"] / 1e9
    kernel_s = r["dev_ms"] / 1e3 / args.steps
    achieved = BYTES_PER_CELL_UPDATE * r["cells"] / kernel_s / 1e9
    log(f"[bench] wall {r['wall_s']:.4f}s  device {r['dev_ms']:.2f}ms  {value:.1f} Gcell/s")

    cpu = None
    if not args.no_cpu:
        log("[bench] cpu baseline")
        cpu = cpu_baseline()
        log(f"[bench] cpu {cpu['value']:.4f} Gcell/s")

    out = {
        "metric": "Gcell-updates/s (16384^2 all-active world)",
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["wall_s"] * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"all-active striped world {args.side}x{args.side}, reactions off, "
                               f"temperature on, {args.steps} ticks",
                   "cells": r["cells"], "cell_iter": 3, "chunk": "64x64", "parallelism": "none (1 GPU)"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_tick_world", "kernel_avg_ms": round(kernel_s * 1e3, 4)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
