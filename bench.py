#!/usr/bin/env python3
"""Throughput benchmark of the rigid-body step path on one MI355X.

Workload: the isolated sort + broadphase pass over 1,048,576 boxes (centres uniform in a
1000 m cube, per-axis half-extents log-uniform in [0.2, 2] m, FP32 boxes, seed 2): base-cell
resolution, grid-entry emission, the 4-digit entry sort, cell-mate and level-link pair
emission, and the canonical pair sort. One step = one such pass; body-steps/s = boxes x passes
per second. Inputs are resident in HBM before the timed region.

Prints ONE JSON line (rank 0). `roofline` covers the kernel with the largest share of device time
in a separately profiled run of the same passes; `cpu_baseline` times the single-threaded C
oracle on the same workload.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "body-steps/s (100k convex pile) + batched vs plain; % HBM peak; bwd/fwd ratio"
BOXES = 1 << 20
SEED = 2


def kernel_bytes(name: str, k: int, entries: int, pairs: int) -> float:
    """Algorithmic HBM bytes of one launch of each broadphase kernel: the minimal payload the
    launch must read and write (entry = key 4 + slot 4 + record 32 bytes; box = 24 bytes of
    min/max; pair = 8 bytes)."""
    key_pay = 8.0
    if name == "entry_sort_pass":        # read key+slot, write key+slot
        return 2.0 * key_pay * entries
    if name == "bp_mates_kernel":        # every sorted key read once; staged pairs written
        return 4.0 * entries + 8.0 * pairs
    if name == "bp_links_kernel":        # box + packed count per collider; pairs written
        return 28.0 * k + 8.0 * pairs
    if name == "bp_emit_kernel":         # box + packed count in; key, slot, record out
        return 28.0 * k + 40.0 * entries
    if name == "bp_count_kernel":        # box in, packed count out
        return 28.0 * k
    if name == "bp_resolution_kernel":   # box in
        return 24.0 * k
    if name == "pair_sort_pass":         # 3 lanes in and out
        return 24.0 * pairs
    if name == "pair_sort_histogram":
        return 8.0 * pairs
    return 0.0


def run_gpu(steps: int, warmup: int):
    import dynamis_amd
    from dynamis_amd import abi
    from dynamis_amd.broadphase import Broadphase
    from dynamis_amd.device import Event, Stream, profile_collect, profile_enable
    from dynamis_amd.scenes import uniform_boxes

    if dynamis_amd.device_count() < 1:
        raise SystemExit("bench.py needs a HIP device")
    dynamis_amd._native.call("dyn_set_device", 0)
    boxes = uniform_boxes(BOXES, 1000.0, 0.2, 2.0, SEED)
    stream = Stream()
    bp = Broadphase(BOXES, stream=stream)
    bp.upload(boxes)
    for _ in range(warmup):
        bp.run()
    stream.synchronize()
    res = bp.read()
    entries = abi.counter(res.counters, "ENTRIES")
    pairs = abi.counter(res.counters, "PAIRS")

    start, stop = Event(), Event()
    start.record(stream)
    for _ in range(steps):
        bp.run()
    stop.record(stream)
    ms = start.elapsed_ms(stop)
    stream.synchronize()

    # Same passes again with every launch bracketed by events on the bench stream.
    profile_enable(True)
    for _ in range(steps):
        bp.run()
    stream.synchronize()
    profile_enable(False)
    stages = profile_collect()
    return ms, entries, pairs, stages


def run_cpu(passes: int):
    from dynamis_amd.scenes import uniform_boxes
    from tests.oracle_binding import load

    oracle = load()
    boxes = uniform_boxes(BOXES, 1000.0, 0.2, 2.0, SEED)
    t0 = time.perf_counter()
    for _ in range(passes):
        oracle.broadphase(boxes)
    return time.perf_counter() - t0


def main() -> int:
    parser = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=50)
    parser.add_argument("--warmup", type=int, default=10)
    parser.add_argument("--cpu-passes", type=int, default=3,
                        help="oracle passes timed for the CPU baseline (0 skips it)")
    args = parser.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(independent worlds would shard by world range; see DESIGN.md)", file=sys.stderr)
        return 2
    steps, warmup = max(args.steps, 1), max(args.warmup, 0)
    ms, entries, pairs, stages = run_gpu(steps, warmup)
    value = BOXES * steps / (ms / 1000.0)

    table = {}
    for name, (launches, total_ms) in stages.items():
        avg_us = 1000.0 * total_ms / max(launches, 1)
        nbytes = kernel_bytes(name, BOXES, entries, pairs)
        table[name] = {"launches": launches, "avg_us": round(avg_us, 3),
                       "share": round(total_ms / max(sum(t for _, t in stages.values()), 1e-9), 4),
                       "bytes": nbytes,
                       "gbs": round(nbytes / (avg_us * 1e-6) / 1e9, 1) if avg_us > 0 else 0.0}
    dominant = max((n for n in table if table[n]["bytes"] > 0),
                   key=lambda n: stages[n][1], default=None)
    roofline = None
    if dominant:
        d = table[dominant]
        achieved = d["gbs"]
        roofline = {"bound": "hbm", "kernel": dominant, "achieved": achieved,
                    "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                    "avg_us": d["avg_us"], "bytes_per_launch": d["bytes"]}

    cpu = None
    if args.cpu_passes > 0:
        secs = run_cpu(args.cpu_passes)
        cpu = {"value": BOXES * args.cpu_passes / secs, "unit": "body-steps/s", "cores": 1,
               "kind": "port",
               "sample": f"{args.cpu_passes} full passes of the same 2^20-box workload through "
                         "the single-threaded C oracle (emission, sort, pairs, pair sort)"}

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": ms / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": "isolated sort + broadphase pass over 1,048,576 boxes "
                               "(centres U[0,1000)^3 m, half-extents log-U[0.2,2] m)",
                   "boxes": BOXES, "entries": entries, "pairs": pairs, "seed": SEED,
                   "step": "resolution + entry emission + 4-digit entry sort + pair emission + "
                           "canonical pair sort"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "stages": table,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
