"""Autotuner for the library's launch parameters, per size bucket.

    python -m ptilopsis_amd.tune [--out PATH]   (on the GPU; default ptilopsis_amd/tuning.json)

Knobs (include/ptl.h, ptl_tune_set): the block size of each stage's per-node
and per-word kernels (64, 128 or 256 threads) and the register-allocation
walks' ring (4 or 8 entries).  Buckets: a compile's node count below 64Ki, 1Mi,
8Mi, or larger.  For each bucket a representative workload is compiled with
each value of each knob in turn (the others at their current choice), and the
value with the smallest device time of the stage the knob affects is kept
(the default stays unless another is at least 2% faster).  The output never
depends on the knobs; the tests check that under non-default values.

apply_defaults() loads tuning.json into the library's process-wide defaults;
Context and Server call it before creating their contexts.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
TUNING_PATH = os.path.join(HERE, "tuning.json")

BUCKETS = ("< 64Ki nodes", "< 1Mi nodes", "< 8Mi nodes", ">= 8Mi nodes")
# knob -> (allowed values, default, index of the stage it affects in stage_ns)
KNOBS = {
    "block.preprocess": ((64, 128, 256), 256, 0),
    "block.isn_cnt": ((64, 128, 256), 256, 1),
    "block.isn_gen": ((64, 128, 256), 256, 2),
    "block.optimize": ((64, 128, 256), 256, 3),
    "block.emit": ((64, 128, 256), 256, 4),
    "block.fix_jumps": ((64, 128, 256), 256, 5),
    "walk_ring": ((4, 8), 4, 4),
    # programs of at most 64 functions allocated by one wave each (1) or by the
    # walk and the lockstep (0)
    "small_programs": ((0, 1), 0, 4),
}

_applied = False


def load() -> dict | None:
    if not os.path.exists(TUNING_PATH):
        return None
    with open(TUNING_PATH) as f:
        return json.load(f)


def apply_defaults(lib) -> None:
    """Sets the library's process-wide defaults from tuning.json (once)."""
    global _applied
    if _applied:
        return
    _applied = True
    t = load()
    if not t:
        return
    for knob, vals in t.get("values", {}).items():
        if knob not in KNOBS:
            continue
        for b, v in enumerate(vals[:len(BUCKETS)]):
            if v in KNOBS[knob][0]:
                lib.ptl_tune_set(None, knob.encode(), b, int(v))


def _measure(ctx, slot: int, stage: int, reps: int = 3) -> float:
    """Smallest device time (ms) of stage `stage` over reps compiles (after one warm compile)."""
    ctx.compile_resident_batch(slot, copy_out=False, results=False)
    best = float("inf")
    for _ in range(reps):
        r = ctx.compile_resident_batch(slot, copy_out=False)[0]
        best = min(best, r["stage_ns"][stage] / 1e6)
    return best


def main() -> int:
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=TUNING_PATH, help="where to write the table (default: the package's tuning.json)")
    args = ap.parse_args()
    from ptilopsis_amd import workloads as W
    from ptilopsis_amd.generator import Context
    t0 = time.time()
    work = {  # bucket -> (name, trees), representative of the configurations in it
        0: ("single_50k", [W.single_50k()]),
        1: ("program_500k", [W.program_of_size(0x5EED0010, 500_000)]),
        2: ("program_4m", [W.sized_program(0x5EED0011, 4_000_000, 5000)]),
        3: ("large_16m", [W.large_16m()]),
    }
    ring_work_3 = ("stress_1024", W.stress_1024())  # the walks bound this batch; in the one 17M-node program they hide behind the lockstep
    print(f"workloads generated in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    ctx = Context(0)
    values = {k: [v[1]] * len(BUCKETS) for k, v in KNOBS.items()}
    measured = {}
    for b in range(len(BUCKETS)):
        name, trees = work[b]
        ctx.upload_batch(1, trees)
        measured[BUCKETS[b]] = {}
        for knob, (allowed, default, stage) in KNOBS.items():
            slot = 1
            if knob == "walk_ring" and b == 3:
                ctx.upload_batch(2, ring_work_3[1])
                slot, wname = 2, ring_work_3[0]
            else:
                wname = name
            times = {}
            for v in allowed:
                ctx.lib.ptl_tune_set(ctx.h, knob.encode(), b, v)
                times[v] = _measure(ctx, slot, stage)
            best = min(times, key=times.get)
            choice = best if times[best] < 0.98 * times[default] else default
            values[knob][b] = choice
            ctx.lib.ptl_tune_set(ctx.h, knob.encode(), b, choice)
            measured[BUCKETS[b]][knob] = {"workload": wname, "stage_ms": {str(k): round(v, 4) for k, v in times.items()},
                                          "choice": choice}
            print(f"bucket {b} {knob:18s} {wname:12s} " +
                  " ".join(f"{k}:{v:.3f}ms" for k, v in times.items()) + f" -> {choice}", file=sys.stderr, flush=True)
    ctx.close()
    import torch
    out = {"device": torch.cuda.get_device_name(0) if torch.cuda.is_available() else "",
           "buckets": list(BUCKETS), "values": values, "measured": measured}
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(values), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
