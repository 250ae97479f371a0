"""Autotuner for the engine's launch shapes.

For each filter stage and each sample size (8-bit frames, 10/12-bit frames)
it measures the stage's kernel time (HIP events on the engine's stream,
batched persistent path) over a sweep of
  - the kernel's register budget: the compiled variant that allows at least
    6, 7 or 8 waves per SIMD, or the compiler's own choice,
  - persistent workgroups per CU (1 .. as many as fit),
  - work items claimed per atomic on the stage counter (1, 2, 4, 8, or a
    static grid-stride walk without atomics),
  - superres rows per work item (4, 8, 16, 32),
  - loop-restoration tile width (32, 64),
  - sub-batches on streams of their own (1-4; timed on whole graph-replayed
    steps, since concurrent stages cannot be timed one by one),
one knob at a time, keeps the fastest (a candidate must beat the current
choice by more than 1.5% to replace it), and checks that the outputs of the
chosen shape are byte-identical to those of the default shape (no knob
changes arithmetic).  The result is a small JSON table per device
(tuning_<arch>.json beside this file) that Engine applies when it exists.

    python -m oxav1.autotune --write     # on the GPU: sweep and save the table
"""
import argparse
import json
import os
import time

import numpy as np

from . import abi

STAGES = ("lf", "cdef", "sr", "lr", "fg")
KINDS = ("u8", "u16")
TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_gfx950.json")


def to_struct(table):
    t = abi.Tuning()
    for s, name in enumerate(STAGES):
        for k, kind in enumerate(KINDS):
            t.blocks_per_cu[s][k] = int(table.get("blocks_per_cu", {}).get(kind, {}).get(name, 0))
            t.waves_per_eu[s][k] = int(table.get("waves_per_eu", {}).get(kind, {}).get(name, 0))
            t.items_per_claim[s][k] = int(table.get("items_per_claim", {}).get(kind, {}).get(name, 0))
    for k, kind in enumerate(KINDS):
        t.sr_rows_per_item[k] = int(table.get("sr_rows_per_item", {}).get(kind, 0))
        t.lr_tile_w[k] = int(table.get("lr_tile_w", {}).get(kind, 0))
        t.sub_batches[k] = int(table.get("sub_batches", {}).get(kind, 0))
    return t


def load_table(path=TABLE):
    """The saved table as a dict, or None when there is none."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def _stage_ms(eng, stage, steps):
    eng.set_timing(True)
    for _ in range(steps):
        eng.run(1)
    eng.sync()
    ms = eng.stage_ms(stage)[0]
    eng.set_timing(False)
    return ms


def _step_ms(eng, steps):
    """Whole batched step (hipGraph replay), ms."""
    for _ in range(2):
        eng.run(2)
    eng.sync()
    t = time.perf_counter()
    for _ in range(steps * 2):
        eng.run(2)
    eng.sync()
    return (time.perf_counter() - t) / (steps * 2) * 1e3


def _outputs(eng, dfs, mode=1):
    for df in dfs:
        df.clear_outputs()
    eng.run(mode)
    eng.sync()
    return [df.result() for df in dfs]


def _same(a, b):
    return all(np.array_equal(x, y) for fa, fb in zip(a, b) for x, y in zip(fa[0] + fa[1], fb[0] + fb[1]))


MARGIN = 0.985
WAVES = (0, 6, 7, 8)
CLAIM_STATIC = 255


def _pick(res, current):
    """The fastest candidate, unless it beats `current` by less than the margin."""
    best = min(res, key=res.get)
    return best if res[best] < res[current] * MARGIN else current


def tune_kind(dfs, kind, steps=6, log=print):
    """Sweep the knobs for one sample size on these device frames; returns
    (waves_per_eu {stage: w}, blocks_per_cu {stage: n}, sr_rows, lr_tile_w,
    timings)."""
    from .engine import Engine
    k = KINDS.index(kind)
    eng = Engine(max_frames=len(dfs), tuning=None)
    try:
        t = abi.Tuning()
        eng.set_tuning(t)
        eng.prepare(dfs)
        eng.run(1)
        eng.sync()
        ref = _outputs(eng, dfs)
        _, cus, fit = eng.tuning()
        active = [s for s in STAGES if eng.stage_items(s)]
        timings = {}
        best_w, best_bpc, best_claim = {}, {}, {}

        def measure(stage):
            eng.set_tuning(t)
            eng.prepare(dfs)
            return _stage_ms(eng, stage, steps)

        for si, s in enumerate(STAGES):
            if s not in active:
                continue
            res = {}
            for w in WAVES:
                t.waves_per_eu[si][k] = w
                t.blocks_per_cu[si][k] = 0
                res[w] = measure(s)
            w = _pick(res, 0)
            t.waves_per_eu[si][k] = w
            best_w[s] = w
            timings[f"{s}_waves_per_eu"] = {str(c): round(v, 4) for c, v in res.items()}
            log(f"  {kind} {s}: waves/SIMD variant {res} -> {w}")
            nfit = fit[si][WAVES.index(w)]
            res = {0: res[w]}
            for c in sorted({c for c in (1, 2, 3, 4, 6, 8) if c < nfit}):
                t.blocks_per_cu[si][k] = c
                res[c] = measure(s)
            c = _pick(res, 0)
            t.blocks_per_cu[si][k] = c
            best_bpc[s] = c
            timings[f"{s}_blocks_per_cu"] = {str(c): round(v, 4) for c, v in res.items()}
            log(f"  {kind} {s}: blocks/CU (0 = all {nfit} that fit) {res} -> {c}")
            res = {1: res[c]}
            for cl in (2, 4, 8, CLAIM_STATIC):
                t.items_per_claim[si][k] = cl
                res[cl] = measure(s)
            cl = _pick(res, 1)
            t.items_per_claim[si][k] = cl
            best_claim[s] = cl
            timings[f"{s}_items_per_claim"] = {str(c): round(v, 4) for c, v in res.items()}
            log(f"  {kind} {s}: items per claim (255 = static) {res} -> {cl}")
        sr_best = 0
        if "sr" in active:
            res = {}
            for r in (8, 4, 16, 32):
                t.sr_rows_per_item[k] = r
                res[r] = measure("sr")
            sr_best = _pick(res, 8)
            t.sr_rows_per_item[k] = sr_best
            timings["sr_rows_per_item"] = {str(c): round(v, 4) for c, v in res.items()}
            log(f"  {kind} sr rows/item {res} -> {sr_best}")
        lr_best = 0
        if "lr" in active:
            res = {}
            for w in (64, 32):
                t.lr_tile_w[k] = w
                res[w] = measure("lr")
            lr_best = _pick(res, 64)
            t.lr_tile_w[k] = lr_best
            timings["lr_tile_w"] = {str(c): round(v, 4) for c, v in res.items()}
            log(f"  {kind} lr tile width {res} -> {lr_best}")
        # sub-batches overlap stages of different frames: timed on whole
        # batched steps (graph replay), since per-stage events cannot
        # separate concurrent stages
        res = {}
        for nsub in (1, 2, 3, 4):
            t.sub_batches[k] = nsub
            eng.set_tuning(t)
            eng.prepare(dfs)
            res[nsub] = _step_ms(eng, steps)
        sub_best = _pick(res, 1)
        t.sub_batches[k] = sub_best
        timings["sub_batches_step_ms"] = {str(c): round(v, 4) for c, v in res.items()}
        log(f"  {kind} sub-batches (whole step ms) {res} -> {sub_best}")
        eng.set_tuning(t)
        eng.prepare(dfs)
        if not _same(_outputs(eng, dfs, 2), ref):
            raise RuntimeError(f"{kind}: tuned launch shape changed the output")
        return best_w, best_bpc, best_claim, sr_best, lr_best, sub_best, timings
    finally:
        eng.close()


def workloads(frames_per_kind=4):
    """Device frames the table is tuned on: 3840x2160 4:2:0 with superres
    12/8, 128x128 superblocks and units, loop restoration and film grain
    with overlap, at 10-bit (u16) and 8-bit (u8)."""
    from .engine import DeviceFrame
    from .frames import Frame, workload_seeds
    lib = abi.load()
    out = {}
    for kind, bd in (("u16", 10), ("u8", 8)):
        dfs = []
        for s in workload_seeds(2, frames_per_kind):
            c = abi.GenConfig()
            abi.check(lib.oxav1_gen_default_config(2, s, c), "gen_default_config")
            c.bit_depth = bd
            c.noise_amp = 32 if bd == 10 else 8
            dfs.append(DeviceFrame(Frame.generate(c, s)))
        out[kind] = dfs
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--write", action="store_true", help=f"save the table to {TABLE}")
    ap.add_argument("--out", default=None, help="save the table here instead")
    ap.add_argument("--frames", type=int, default=8)
    ap.add_argument("--steps", type=int, default=6)
    args = ap.parse_args()
    abi.check(abi.load().oxav1_pf_init(), "init")
    from .engine import device_info
    t0 = time.time()
    table = {"device": {"cus": device_info()[0]}, "waves_per_eu": {}, "blocks_per_cu": {},
             "items_per_claim": {}, "sr_rows_per_item": {}, "lr_tile_w": {}, "sub_batches": {},
             "timings_ms": {}}
    for kind, dfs in workloads(args.frames).items():
        wv, bpc, claim, sr, lr, sub, tm = tune_kind(dfs, kind, args.steps)
        table["items_per_claim"][kind] = claim
        table["sub_batches"][kind] = sub
        table["waves_per_eu"][kind] = wv
        table["blocks_per_cu"][kind] = bpc
        table["sr_rows_per_item"][kind] = sr
        table["lr_tile_w"][kind] = lr
        table["timings_ms"][kind] = tm
        for df in dfs:
            df.free()
    table["workload"] = (f"{args.frames} x 3840x2160 4:2:0 frames per sample size (coded 2560 wide, "
                         f"superres, loop restoration, film grain), batched path, {args.steps} timed runs "
                         f"per point")
    table["tune_s"] = round(time.time() - t0, 1)
    text = json.dumps(table, indent=1, sort_keys=True)
    print(text)
    path = args.out or (TABLE if args.write else None)
    if path:
        with open(path, "w") as f:
            f.write(text + "\n")


if __name__ == "__main__":
    main()
