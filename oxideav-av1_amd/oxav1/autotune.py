"""Autotuner for the engine's launch shapes.

For each filter stage and each sample size (8-bit frames, 10/12-bit frames)
it measures the stage's kernel time (HIP events on the engine's stream,
batched persistent path) over a sweep of
  - persistent workgroups per CU (1 .. as many as fit),
  - superres rows per work item (4, 8, 16, 32),
  - loop-restoration tile width (32, 64),
one knob at a time, keeps the fastest, and checks that the outputs of the
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
    for k, kind in enumerate(KINDS):
        t.sr_rows_per_item[k] = int(table.get("sr_rows_per_item", {}).get(kind, 0))
        t.lr_tile_w[k] = int(table.get("lr_tile_w", {}).get(kind, 0))
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


def _outputs(eng, dfs):
    for df in dfs:
        df.clear_outputs()
    eng.run(1)
    eng.sync()
    return [df.result() for df in dfs]


def _same(a, b):
    return all(np.array_equal(x, y) for fa, fb in zip(a, b) for x, y in zip(fa[0] + fa[1], fb[0] + fb[1]))


def tune_kind(dfs, kind, steps=4, log=print):
    """Sweep the knobs for one sample size on these device frames; returns
    (blocks_per_cu {stage: n}, sr_rows, lr_tile_w, timings)."""
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
        best_bpc = {}
        for si, s in enumerate(STAGES):
            if s not in active:
                continue
            cands = sorted({c for c in (1, 2, 3, 4, 6, 8, fit[si]) if 1 <= c <= fit[si]})
            res = {}
            for c in cands:
                t.blocks_per_cu[si][k] = c
                eng.set_tuning(t)
                eng.prepare(dfs)
                res[c] = _stage_ms(eng, s, steps)
            best = min(res, key=res.get)
            t.blocks_per_cu[si][k] = best
            best_bpc[s] = best
            timings[f"{s}_blocks_per_cu"] = {str(c): round(v, 4) for c, v in res.items()}
            log(f"  {kind} {s}: blocks/CU {res} -> {best}")
        sr_best = 0
        if "sr" in active:
            res = {}
            for r in (4, 8, 16, 32):
                t.sr_rows_per_item[k] = r
                eng.set_tuning(t)
                eng.prepare(dfs)
                res[r] = _stage_ms(eng, "sr", steps)
            sr_best = min(res, key=res.get)
            t.sr_rows_per_item[k] = sr_best
            timings["sr_rows_per_item"] = {str(c): round(v, 4) for c, v in res.items()}
            log(f"  {kind} sr rows/item {res} -> {sr_best}")
        lr_best = 0
        if "lr" in active:
            res = {}
            for w in (32, 64):
                t.lr_tile_w[k] = w
                eng.set_tuning(t)
                eng.prepare(dfs)
                res[w] = _stage_ms(eng, "lr", steps)
            lr_best = min(res, key=res.get)
            t.lr_tile_w[k] = lr_best
            timings["lr_tile_w"] = {str(c): round(v, 4) for c, v in res.items()}
            log(f"  {kind} lr tile width {res} -> {lr_best}")
        eng.set_tuning(t)
        eng.prepare(dfs)
        if not _same(_outputs(eng, dfs), ref):
            raise RuntimeError(f"{kind}: tuned launch shape changed the output")
        return best_bpc, sr_best, lr_best, timings
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
    ap.add_argument("--frames", type=int, default=4)
    ap.add_argument("--steps", type=int, default=4)
    args = ap.parse_args()
    abi.check(abi.load().oxav1_pf_init(), "init")
    from .engine import device_info
    t0 = time.time()
    table = {"device": {"cus": device_info()[0]}, "blocks_per_cu": {}, "sr_rows_per_item": {},
             "lr_tile_w": {}, "timings_ms": {}}
    for kind, dfs in workloads(args.frames).items():
        bpc, sr, lr, tm = tune_kind(dfs, kind, args.steps)
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
