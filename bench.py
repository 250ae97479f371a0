#!/usr/bin/env python3
"""Benchmark: byte-exact post-filter Mpixel/s of the AV1 filter chain on one
MI355X, batched vs plain, with per-frame latency.

Workload (a step): one batch of `--frames` seeded synthetic 3840x2160
10-bit 4:2:0 frames (coded 2560 wide, superres denominator 12, 128x128
superblocks, loop restoration with 128x128 Wiener/self-guided units, film
grain with overlap) pushed through deblock -> CDEF -> superres -> loop
restoration -> film grain.  Inputs (pre-filter planes, grids, parameters)
are resident in HBM before the timed region.

`--cfg 3 --frames 64` runs the mixed batch instead: 64 frames in four
blocks of 640x360, 1280x720, 1920x1080 and 3840x2160, 8/10-bit and
4:2:0/4:2:2/4:4:4 drawn per frame.

value  = output Mpixel/s of the batched path (one persistent launch per
         stage over the whole batch, replayed from a hipGraph).
plain  = the same frames, one launch per stage per frame.
Byte exactness: the batched output equals the plain output for every frame,
and the first frames are compared with the CPU oracle, which is also the
timed CPU baseline (single thread, literal restatement of the reference).
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "oxideav-av1_amd"))

METRIC = "byte-exact post-filter Mpixel/s at 4K 10-bit, batched vs plain; p99 latency"
HBM_PEAK_GBS = 8000.0


def describe(frames):
    """Short description of a list of frames: sizes, depths, subsampling."""
    from collections import Counter
    c = Counter()
    for f in frames:
        d = f.desc
        ss = "mono" if d.num_planes == 1 else {(1, 1): "4:2:0", (1, 0): "4:2:2", (0, 0): "4:4:4"}[
            (d.subsampling_x, d.subsampling_y)]
        sr = f" coded {d.frame_width} wide" if d.upscaled_width > d.frame_width else ""
        c[f"{d.upscaled_width}x{d.frame_height}{sr} {d.bit_depth}-bit {ss}"] += 1
    return ", ".join(f"{n} x {k}" for k, n in c.items())


def stages_on(frames):
    names = {"lf": "deblock", "cdef": "CDEF", "sr": "superres", "lr": "loop restoration", "fg": "film grain"}
    return "+".join(names[k] for k in stage_keys(frames))


def host_info():
    """Host CPU facts for the CPU baseline: model, logical CPUs, usable cores."""
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return {"cpu_model": model, "nproc": os.cpu_count(), "usable_cores": usable}


def dropin_rate(frame, expect=None):
    """The reference's interface end to end: the frame driver's post-filter
    sequence through the five drop-in calls on host int32 planes (each call
    stages its planes to HBM and back over PCIe), the driver's planes reused
    from frame to frame as a decoder's frame buffers are (the frame's
    reconstruction is copied into the driver's working planes before the
    clock starts).  expect: the engine's (pre-grain, output) planes of the
    same frame, compared with the calls' result.  Returns a dict."""
    import numpy as np

    from oxav1 import dropin
    calls = dropin.gpu()
    bufs = dropin.driver_buffers(frame)
    ref_d, out_d = dropin.frame_driver_chain(calls, frame, bufs=bufs)  # warm
    same = None
    if expect is not None:
        same = all(np.array_equal(a, b) for a, b in zip(ref_d + out_d, expect[0] + expect[1]))
    reps, stage = 3, []
    for _ in range(reps):
        st = []
        dropin.frame_driver_chain(calls, frame, st, bufs=bufs)
        stage.append(st)
    s = float(np.sum(np.mean(stage, axis=0)))
    # the same with every plane freshly allocated per call (first-touch page
    # faults and the driver's copies of the original Python harness)
    fresh = []
    for _ in range(2):
        st = []
        dropin.frame_driver_chain(calls, frame, st)
        fresh.append(sum(st))
    d = frame.desc
    return {"mpx_s": round(d.upscaled_width * d.frame_height / s / 1e6, 2), "ms_per_frame": round(s * 1e3, 2),
            "stage_ms": {k: round(v * 1e3, 2) for k, v in zip(("lf", "cdef", "sr", "lr", "fg"),
                                                              np.mean(stage, axis=0))},
            "ms_per_frame_fresh_planes": round(min(fresh) * 1e3, 2),
            "equals_engine_output": same,
            "what": f"one {d.upscaled_width}x{d.frame_height} frame through the five drop-in calls "
                    f"(int32 host planes, PCIe-inclusive), the reference's own interface, the driver's planes "
                    f"reused across frames"}


def stage_keys(frames):
    """Engine stage names that do work on at least one of the frames."""
    keys = ["lf", "cdef"]
    if any(f.desc.use_superres and f.desc.upscaled_width > f.desc.frame_width for f in frames):
        keys.append("sr")
    if any(f.desc.lr.uses_lr for f in frames):
        keys.append("lr")
    if any(f.desc.fg.apply_grain for f in frames):
        keys.append("fg")
    return keys


def measured_traffic(cfg, nframes, stage):
    """HBM bytes per launch of the stage's kernel for this workload, from the
    committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary (calibrated against a
    copy kernel of known size); None when no summary covers it."""
    import re
    path = os.path.join(ROOT, "profiles", f"traffic_cfg{cfg}_batch{nframes}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        m = re.match(r"^traffic_bytes_per_launch\s+(\w+)\s+(\d+)", line.strip())
        if m and m.group(1) == stage:
            return int(m.group(2))
    return None


def serve(args):
    """Serving workload: args.sessions decode sessions with seeds
    splitmix64(0x0A1F0004 ^ i), each args.session-frames frames arriving at
    args.fps into 8 resident reference slots, run twice:
    - device-resident: 8 pre-filter frames per session uploaded once at
      admission (the filter chain alone);
    - host frames: 2 pre-filter frames per session in page-locked host memory,
      every submission copying its frame to the device on the server's upload
      stream and every result copied back to host memory before the frame is
      done (what a decoder with CPU reconstruction pays).
    Every session's held slots (and, with host frames, the copies handed back)
    are compared at its end with the same frames filtered alone."""
    from oxav1 import abi
    from oxav1.engine import DeviceFrame, Engine
    from oxav1.serving import DecodeSession, Server, expected_outputs, run_serving

    lib = abi.load()
    t0 = time.time()
    seeds = [int(lib.oxav1_gen_splitmix64(0x0A1F0004 ^ i)) for i in range(args.sessions)]

    def one(host):
        sessions = [DecodeSession(sd, n_frames=args.session_frames, fps=args.fps, pool=2 if host else 8,
                                  slots=8, host=host) for sd in seeds]
        eng = Engine(max_frames=1)
        expected = expected_outputs(sessions, eng)
        eng.close()
        largest = max((s.desc for s in sessions), key=lambda d: d.upscaled_width * d.frame_height * d.bit_depth)
        srv = Server(lanes=args.lanes, max_batch=args.max_batch, largest=largest)
        try:
            res = run_serving(srv, sessions, expected, max_seconds=600)
        finally:
            srv.close()
        return sessions, res

    sessions, res = one(False)
    # alone latency of one frame of each shape (plain path, enqueue to completion)
    eng = Engine(max_frames=1)
    alone_ms = {}
    for s in sessions:
        key = f"{s.desc.upscaled_width}x{s.desc.frame_height} {s.desc.bit_depth}-bit"
        if key in alone_ms:
            continue
        df = DeviceFrame(s.frames[0])
        eng.prepare([df])
        ts = []
        for i in range(13):
            t = time.perf_counter()
            eng.run(0)
            eng.sync()
            if i >= 3:
                ts.append((time.perf_counter() - t) * 1e3)
        alone_ms[key] = round(statistics.median(ts), 3)
        df.free()
    eng.close()
    hsessions, hres = one(True)
    setup_s = time.time() - t0
    st = res["stats"]
    mix = describe_sessions(sessions)

    def lat(r, what):
        return {"p50": round(r["p50_ms"], 3), "p99": round(r["p99_ms"], 3), "max": round(r["max_ms"], 3),
                "queue_p99": round(r["queue_p99_ms"], 3),
                "launch_to_done_p50": round(r["launch_to_done_p50_ms"], 3),
                "launch_to_done_p99": round(r["launch_to_done_p99_ms"], 3), "what": what}

    return {
        "metric": "serving: sustained post-filter Mpixel/s with p50/p99 per-frame latency, byte-exact",
        "value": round(res["mpx_s"], 1),
        "unit": "Mpixel/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "data": "synthetic",
        "config": {"workload": f"{args.sessions} decode sessions ({mix}), Poisson arrivals at {args.fps:g} fps "
                               f"each, {args.session_frames} frames per session, 8 resident reference slots "
                               f"each; deblock+CDEF+loop restoration",
                   "lanes": args.lanes, "max_batch": args.max_batch},
        "latency_ms": lat(res, "device-resident frames: arrival to completion seen by the serving loop"),
        "host_frames": {
            "mpx_s": round(hres["mpx_s"], 1),
            "latency_ms": lat(hres, "frames in host memory: arrival (submission) to the result back in host "
                                    "memory, copies both ways included"),
            "frames": hres["frames"],
            "byte_exact": {"frames_compared_with_alone": hres["verified_frames"],
                           "mismatches": len(hres["mismatches"])},
            "batching": {"mean_frames_per_launch": round(hres["mean_batch"], 2), "max": hres["max_batch"]},
        },
        "alone_latency_ms": alone_ms,
        "frames": res["frames"],
        "span_s": round(res["span_s"], 3),
        "batching": {"mean_frames_per_launch": round(res["mean_batch"], 2), "max": res["max_batch"],
                     "launches": st["launches"]},
        "byte_exact": {"frames_compared_with_alone": res["verified_frames"],
                       "mismatches": len(res["mismatches"])},
        "hbm": {"arena_gb": round(st["arena_bytes"] / 1e9, 2), "lane_gb": round(st["lane_bytes"] / 1e9, 2),
                "peak_sessions_open": res["peak_sessions_open"]},
        "setup_s": round(setup_s, 1),
    }


def side_paths(reps=20):
    """Latency of the two side paths on the metric's frame shape."""
    import ctypes as C

    import numpy as np

    from oxav1 import Frame, abi
    from oxav1.engine import DevBuffer
    from oxav1.serving import DecodeSession, Server

    lib = abi.load()
    f = Frame.generate(2, 0x0A1F0002)
    d = f.desc
    # walker grids -> packed words and transform sizes
    walker = dict(is_inter=(f.is_intra == 0), skip=f.skip, ref_frame=f.ref_frame, mode=f.mode,
                  segment_id=f.segment_id, mi_size=f.mi_size, tx_size=f.lf_tx_size[0],
                  inter_tx_size=f.lf_tx_size[0])
    keep = {k: DevBuffer(v.shape, np.uint8, np.ascontiguousarray(v, dtype=np.uint8)) for k, v in walker.items()}
    g = abi.WalkerGrids(**{k: b.ptr for k, b in keep.items()})
    lf_mi = DevBuffer((d.mi_rows, d.mi_cols), np.uint32)
    tx = [DevBuffer(f.lf_tx_size[p].shape, np.uint8) for p in range(3)]
    ptrs = (C.c_void_p * 3)(*[t.ptr for t in tx])

    def build():
        abi.check(lib.oxav1_pf_build_grids(C.byref(g), d.mi_rows, d.mi_cols, d.num_planes, d.subsampling_x,
                                           d.subsampling_y, lf_mi.ptr, ptrs, None), "build_grids")
    build()
    exact = np.array_equal(lf_mi.download(), f.lf_mi) and all(
        np.array_equal(tx[p].download(), f.lf_tx_size[p]) for p in range(3))
    t = time.perf_counter()
    for _ in range(reps):
        build()
    lf_mi.download()  # synchronises the default stream
    grid_ms = (time.perf_counter() - t) / reps * 1e3
    for b in list(keep.values()) + [lf_mi] + tx:
        b.free()
    # show_existing_frame on a resident slot
    cfg = abi.GenConfig()
    abi.check(lib.oxav1_gen_default_config(2, 0x0A1F0002, C.byref(cfg)), "config")
    sess = DecodeSession(0x0A1F0002, n_frames=1, pool=1, slots=1, gen_config=cfg)
    srv = Server(lanes=1, max_batch=1)
    try:
        assert sess.admit(srv)
        desc, dev, slot = sess.frame_job(0)
        srv.submit(sess.sid, desc, dev, slot, srv.now_us())
        while not srv.poll():
            pass
        lat = []
        for _ in range(reps):
            t0 = srv.now_us()
            srv.submit_existing(sess.sid, 0, sess.desc.fg, t0)
            done = []
            while not done:
                done = srv.poll()
            lat.append((done[0].done_us - t0) * 1e-3)
    finally:
        srv.close()
    return {"what": "side paths on one 3840x2160 10-bit 4:2:0 frame",
            "build_grids_ms": round(grid_ms, 4), "build_grids_exact": bool(exact),
            "mi_cells": int(d.mi_rows * d.mi_cols),
            "show_existing_ms": {"p50": round(statistics.median(lat), 3), "max": round(max(lat), 3)},
            "show_existing_what": "film grain from the slot's parameters on the resident pre-grain frame, "
                                  "submit to completion seen by the serving loop"}


def elect_bench(reps=5, cpu=True):
    """The encoder's CDEF strength election (encoder/cdef_elect.rs) on one
    frame: the whole election call (host planes in, plan out), one launch of
    the per-unit SSD kernel over the first sweep's candidates on resident
    planes, and the C oracle's election on the 1080p frame for comparison."""
    import ctypes as C

    import numpy as np

    from oxav1 import abi
    from oxav1.elect import elect_cdef, synthetic_input
    from oxav1.engine import DevBuffer

    lib = abi.load()
    out = []
    for w, h, bd in ((1920, 1080, 10), (3840, 2160, 10)):
        inp = synthetic_input(0xE1EC7 + w, w, h, bd, 1, 1, lambda_=600 * 16, max_bits=3)
        r = elect_cdef(inp)
        lat = []
        for _ in range(reps):
            t = time.perf_counter()
            elect_cdef(inp)
            lat.append((time.perf_counter() - t) * 1e3)
        # the SSD kernel alone: the zero rows and the coarse and
        # secondary-only grids of both dampings and plane sets (78 candidates)
        cands = [abi.CdefCand(3, 0, 0, 0), abi.CdefCand(3, 1, 0, 0)]
        for d in (3, 5):
            for ps in (0, 1):
                cands += [abi.CdefCand(d, ps, p, s) for p in (1, 2, 3, 4, 6, 9, 12, 15) for s in (0, 2)]
                cands += [abi.CdefCand(d, ps, 0, s) for s in (1, 2, 4)]
        f = abi.CdefSsdFrame()
        f.mi_rows, f.mi_cols = inp.mi_rows, inp.mi_cols
        keep = [DevBuffer(inp.cdef_idx.shape, np.int8, inp.cdef_idx), DevBuffer(inp.skip.shape, np.uint8, inp.skip)]
        f.cdef_idx, f.skip = keep[0].ptr, keep[1].ptr
        for p in range(3):
            keep += [DevBuffer(inp.recon[p].shape, np.uint16, inp.recon[p]),
                     DevBuffer(inp.input[p].shape, np.uint16, inp.input[p])]
            f.recon[p], f.input[p] = keep[-2].ptr, keep[-1].ptr
            f.rows[p], f.cols[p] = inp.recon[p].shape
            f.recon_pitch[p] = f.input_pitch[p] = inp.recon[p].shape[1]
        f.bit_depth, f.subsampling_x, f.subsampling_y, f.num_planes, f.sample_bytes = bd, 1, 1, 3, 2
        ssd = DevBuffer((len(cands), inp.n_units), np.uint64)
        keep.append(ssd)
        f.ssd = ssd.ptr
        ca = (abi.CdefCand * len(cands))(*cands)
        abi.check(lib.oxav1_pf_cdef_unit_ssd(C.byref(f), 1, ca, len(cands), None), "cdef_unit_ssd")
        kl = []
        for _ in range(reps):
            t = time.perf_counter()
            abi.check(lib.oxav1_pf_cdef_unit_ssd(C.byref(f), 1, ca, len(cands), None), "cdef_unit_ssd")
            kl.append((time.perf_counter() - t) * 1e3)
        for b in keep:
            b.free()
        luma, chroma = w * h, 2 * ((w + 1) // 2) * ((h + 1) // 2)
        samples = sum(luma if c.plane_set == 0 else chroma for c in cands)
        k_ms = statistics.median(kl)
        row = {"frame": f"{w}x{h} {bd}-bit 4:2:0", "units": inp.n_units,
               "elect_ms": {"p50": round(statistics.median(lat), 2), "min": round(min(lat), 2)},
               "ssd_launch_ms": round(k_ms, 3), "ssd_launch_candidates": len(cands),
               "filtered_gsamples_per_s": round(samples / (k_ms * 1e-3) / 1e9, 1),
               "plan": None if r is None else {"cdef_bits": r.best.params["cdef_bits"],
                                               "damping": r.best.params["cdef_damping"],
                                               "d": r.best.d, "base_d": r.base_d}}
        if cpu and w == 1920:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import oracle
            t = time.perf_counter()
            want = elect_cdef(inp, oracle.load(), "oo_cdef_elect")
            cpu_s = time.perf_counter() - t
            same = (want is None and r is None) or (
                want is not None and r is not None and want.best.params == r.best.params and want.best.d == r.best.d
                and np.array_equal(want.best.unit_idx, r.best.unit_idx) and want.base_d == r.base_d)
            row["cpu_oracle"] = {"s": round(cpu_s, 2), "cores": 1, "kind": "port",
                                 "speedup": round(cpu_s * 1e3 / statistics.median(lat), 1)}
            row["byte_exact_vs_oracle"] = bool(same)
        out.append(row)
    # superres and film-grain election helpers, host u16 planes in and out
    from oxav1 import Frame
    from oxav1 import elect as E
    helpers = []
    for w, h in ((1920, 1080), (3840, 2160)):
        inp = synthetic_input(0xF1 + w, w, h, 10, 1, 1)
        pl = inp.input
        cw = E.superres_coded_width(w, 12)
        small = E.downscale_width(pl, 1, 1, cw)
        rec = [small[0]] + [np.resize(p, ((h + 1) >> 1, (cw + 1) >> 1)) for p in small[1:]]
        den = E.fg_denoise(pl, 1, 1)
        fg = Frame.generate(2, 0x0A1F0002).desc.fg  # the 4K workload's grain parameters
        ops = {
            "superres_probe": lambda o=None: E.superres_probe(pl[0], 10, o),
            "superres_downscale": lambda o=None: E.downscale_width(pl, 1, 1, cw, o),
            "superres_upscale_recon": lambda o=None: E.upscale_recon(rec, w, (cw + 3) // 4, 10, 1, 1, o),
            "fg_denoise": lambda o=None: E.fg_denoise(pl, 1, 1, o),
            "fg_noise_estimate": lambda o=None: E.fg_noise_estimate(pl, den, 1, 1, 10, o),
            "fg_apply_grain": lambda o=None: E.fg_apply_grain(fg, pl, 10, 1, 1, 2, o),
        }
        row = {"frame": f"{w}x{h} 10-bit 4:2:0", "gpu_ms": {}}
        if cpu and w == 1920:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import oracle
            row["cpu_oracle_ms"], row["equal"] = {}, {}
        for name, fn in ops.items():
            r = fn()
            t = []
            for _ in range(reps):
                t0 = time.perf_counter()
                fn()
                t.append((time.perf_counter() - t0) * 1e3)
            row["gpu_ms"][name] = round(statistics.median(t), 3)
            if "cpu_oracle_ms" in row:
                t0 = time.perf_counter()
                want = fn(oracle.load())
                row["cpu_oracle_ms"][name] = round((time.perf_counter() - t0) * 1e3, 2)
                if isinstance(r, list):
                    row["equal"][name] = all(np.array_equal(a, b) for a, b in zip(r, want))
                else:
                    row["equal"][name] = r == want
        helpers.append(row)
    # loop-restoration election (encoder/lr_elect.rs), host u16 planes in
    lr_rows = []
    for w, h in ((640, 360), (1920, 1080), (3840, 2160)):
        inp = E.synthetic_lr_input(0x1E + w, w, h, 10, 1, 1, lambda_=600 * 16)
        r = E.elect_lr(inp)
        call = E.lr_elect_call(inp)
        lat = []
        for _ in range(reps):
            t0 = time.perf_counter()
            call()
            lat.append((time.perf_counter() - t0) * 1e3)
        row = {"frame": f"{w}x{h} 10-bit 4:2:0", "units": sum(len(u) for u in r.units),
               "elect_ms": {"p50": round(statistics.median(lat), 2), "min": round(min(lat), 2)},
               "elected": r.elected, "frame_restoration_type": list(r.frame_restoration_type),
               "d": r.d, "d_pre": r.d_pre}
        if cpu and w == 640:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import oracle
            t0 = time.perf_counter()
            want = E.elect_lr(inp, oracle.load(), "oo_lr_elect")
            cpu_s = time.perf_counter() - t0
            row["cpu_oracle"] = {"s": round(cpu_s, 2), "cores": 1, "kind": "port",
                                 "speedup": round(cpu_s * 1e3 / statistics.median(lat), 1)}
            row["equal_to_oracle"] = (want.units == r.units and want.cands == r.cands and want.d == r.d
                                      and want.d_pre == r.d_pre and want.elected == r.elected)
        lr_rows.append(row)
    return {"what": "encoder CDEF strength election (sweeps, damping refinement, greedy per-unit sets, "
                    "D*256 + lambda*R256), host u16 planes in, elected plan out",
            "frames": out,
            "lr_what": "encoder loop-restoration election (Wiener and self-guided fits per 64x64 unit, "
                       "exact evaluation of 18 candidates per unit, D*256 + lambda*R256), host u16 planes in, "
                       "plan and per-unit candidate records out",
            "lr_frames": lr_rows,
            "helpers_what": "superres and film-grain election helpers (superres_elect.rs, film_grain_elect.rs), "
                            "host u16 planes in and out (PCIe-inclusive), superres denominator 12",
            "helpers": helpers}


def describe_sessions(sessions):
    from collections import Counter
    c = Counter(f"{s.desc.upscaled_width}x{s.desc.frame_height} {s.desc.bit_depth}-bit" for s in sessions)
    return ", ".join(f"{n} x {k}" for k, n in sorted(c.items()))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--frames", type=int, default=8, help="frames per step (batch)")
    ap.add_argument("--cfg", type=int, default=2, help="workload configuration (2 = 4K 10-bit)")
    ap.add_argument("--cpu-frames", type=int, default=4, help="frames timed on the CPU oracle (about 10 s)")
    ap.add_argument("--cpu-reps", type=int, default=0,
                    help="repetitions of the CPU sample (0: 50 for frames under 1 Mpixel, else 1)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-parallel", type=int, default=1,
                    help="also time the CPU oracle frame-parallel over the host cores (1/0)")
    ap.add_argument("--no-dropin", action="store_true", help="skip the drop-in (PCIe-inclusive) timing")
    ap.add_argument("--profile", action="store_true", help="timed batched steps only (for rocprofv3)")
    ap.add_argument("--serve", action="store_true",
                    help="serving loop instead: concurrent decode sessions (720p/1080p/4K, 8/10-bit 4:2:0), "
                         "Poisson arrivals, continuous batching across lanes")
    ap.add_argument("--sessions", type=int, default=32)
    ap.add_argument("--fps", type=float, default=30.0, help="frame arrival rate per session")
    ap.add_argument("--session-frames", type=int, default=60, help="frames each session decodes")
    ap.add_argument("--lanes", type=int, default=4)
    ap.add_argument("--max-batch", type=int, default=16)
    ap.add_argument("--side", action="store_true",
                    help="time the side paths on one 3840x2160 10-bit frame: side information built on the "
                         "GPU from walker grids, and the show_existing_frame grain output of a resident slot")
    ap.add_argument("--elect", action="store_true",
                    help="time the encoder's CDEF and loop-restoration elections and the superres and film-grain "
                         "election helpers on 1080p and 4K 10-bit frames")
    ap.add_argument("--elect-lr", action="store_true",
                    help="only the loop-restoration election's raw C call at 1080p and 4K (one warm call, "
                         "--steps timed calls each), for profiling")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="run only the plain device copy kernel over this many bytes, --steps times "
                         "(calibrates the FETCH_SIZE / WRITE_SIZE counters against a known traffic)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU; runs on more than one GPU are not part of it "
              "(frames and sessions shard across GPUs with no exchange between them)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from oxav1 import Frame, abi, workload_seeds
    from oxav1.engine import DeviceFrame, Engine, device_info
    from oxav1.traffic import stage_bytes

    lib = abi.load()
    abi.check(lib.oxav1_pf_init(), "init")
    num_cus, _, hbm_total = device_info()

    if args.copy_bytes:
        from oxav1.engine import DevBuffer
        n = (args.copy_bytes + 15) // 16 * 16
        src = DevBuffer((n,), np.uint8)
        dst = DevBuffer((n,), np.uint8)
        src.fill(1)
        t = time.perf_counter()
        abi.check(lib.oxav1_pf_dev_copy_kernel(dst.ptr, src.ptr, n, args.steps), "copy kernel")
        s = (time.perf_counter() - t) / max(args.steps, 1)
        print(json.dumps({"copy_kernel_bytes": n, "launches": args.steps,
                          "read_bytes_per_launch": n, "write_bytes_per_launch": n,
                          "approx_gbs": round(2 * n / s / 1e9, 1)}))
        src.free()
        dst.free()
        return

    if args.serve:
        print(json.dumps(serve(args)))
        return
    if args.side:
        print(json.dumps(side_paths()))
        return
    if args.elect:
        print(json.dumps(elect_bench(cpu=not args.no_cpu_baseline)))
        return
    if args.elect_lr:
        from oxav1 import elect as E
        rows = []
        for w, h in ((1920, 1080), (3840, 2160)):
            inp = E.synthetic_lr_input(0x1E + w, w, h, 10, 1, 1, lambda_=600)
            call = E.lr_elect_call(inp)
            call()
            lat = []
            for _ in range(args.steps):
                t0 = time.perf_counter()
                call()
                lat.append((time.perf_counter() - t0) * 1e3)
            rows.append({"frame": f"{w}x{h} 10-bit 4:2:0", "calls": args.steps,
                         "elect_ms_p50": round(statistics.median(lat), 2)})
        print(json.dumps({"what": "oxav1_pf_lr_elect raw call, host u16 planes in, plan out", "rows": rows}))
        return

    seeds = workload_seeds(args.cfg, args.frames)
    t0 = time.time()
    frames = [Frame.generate(args.cfg, s) for s in seeds]
    dfs = [DeviceFrame(f) for f in frames]
    gen_s = time.time() - t0
    d0 = frames[0].desc
    fpx = [f.desc.upscaled_width * f.desc.frame_height for f in frames]
    total_px = sum(fpx)

    eng = Engine(max_frames=len(dfs))
    eng.prepare(dfs)

    def timed(mode, steps, warmup):
        for _ in range(warmup):
            eng.run(mode)
        eng.sync()
        t = time.perf_counter()
        for _ in range(steps):
            eng.run(mode)
        eng.sync()
        return (time.perf_counter() - t) / steps

    if args.profile:
        ms = timed(2, args.steps, args.warmup) * 1e3
        print(json.dumps({"profile_run": True, "ms_per_step": ms}))
        eng.close()
        return

    # --- batched (persistent launches, hipGraph replay): the headline value ---
    s_batch = timed(2, args.steps, args.warmup)
    # --- plain (one launch per stage per frame) ---
    s_plain = timed(0, max(3, args.steps // 3), 2)
    # --- batched without the graph ---
    s_batch_nograph = timed(1, max(3, args.steps // 3), 2)

    # byte exactness: batched == plain on every frame
    eng.run(2)
    eng.sync()
    batched = [df.result() for df in dfs]
    for df in dfs:
        df.clear_outputs()
    eng.run(0)
    eng.sync()
    plain = [df.result() for df in dfs]
    exact_paths = sum(all(np.array_equal(a, b) for a, b in zip(pb[0] + pb[1], pp[0] + pp[1]))
                      for pb, pp in zip(batched, plain))

    # --- per-stage kernel time (HIP events on the engine stream, batched launches) ---
    eng.set_timing(True)
    for _ in range(max(3, args.steps // 3)):
        eng.run(1)
    eng.sync()
    active = stage_keys(frames)
    stage_ms = {s: eng.stage_ms(s)[0] for s in active}
    eng.set_timing(False)
    sb = {k: sum(stage_bytes(f.desc)[k] for f in frames) for k in stage_ms}
    dom = max(stage_ms, key=lambda k: stage_ms[k])
    achieved = sb[dom] / (stage_ms[dom] * 1e-3) / 1e9
    chain_alg = sum(sb.values())
    chain_gbs = chain_alg / s_batch / 1e9

    # --- per-frame latency: one frame alone, plain path ---
    eng1 = Engine(max_frames=1)
    eng1.prepare([dfs[0]])
    lat = []
    for i in range(max(20, args.steps)):
        t = time.perf_counter()
        eng1.run(0)
        eng1.sync()
        if i >= 3:
            lat.append((time.perf_counter() - t) * 1e3)
    eng1.close()
    lat.sort()
    p50 = statistics.median(lat)
    p99 = lat[min(len(lat) - 1, int(round(0.99 * (len(lat) - 1))))]

    # --- CPU baseline (oracle, single thread) on a bounded sample; also the oracle check ---
    cpu = None
    oracle_exact = None
    if not args.no_cpu_baseline and args.cpu_frames > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        oracle.load()
        n = min(args.cpu_frames, len(frames))
        # evenly spaced over the batch, so a mixed batch's sample covers each of its blocks
        pick = [i * len(frames) // n for i in range(n)]
        reps = args.cpu_reps or (50 if max(fpx[i] for i in pick) < 1_000_000 else 1)
        t = time.perf_counter()
        for _ in range(reps):
            outs = [oracle.chain(frames[i]) for i in pick]
        cpu_s = (time.perf_counter() - t) / reps
        oracle_exact = 0
        for i, (ref_o, out_o, _) in zip(pick, outs):
            ref_g, out_g = plain[i]
            oracle_exact += all(np.array_equal(a, b) for a, b in zip(ref_o + out_o, ref_g + out_g))
        which = f"the first {n}" if pick == list(range(n)) else f"{n} evenly spaced"
        stage_ns = np.mean([o[2] for o in outs], axis=0)
        cpu = {"value": round(sum(fpx[i] for i in pick) / cpu_s / 1e6, 4), "unit": "Mpixel/s", "cores": 1,
               "kind": "port",
               "sample": f"{which} of the {len(frames)} frames of the step ({describe([frames[i] for i in pick])}), "
                         f"full chain through the C oracle, single thread, "
                         + (f"{reps} repetitions, {cpu_s * 1e3:.2f} ms each" if reps > 1 else f"{cpu_s:.1f} s"),
               "stage_ms_per_frame": {k: round(v * 1e-6, 1) for k, v in zip(("lf", "cdef", "sr", "lr", "fg"),
                                                                          stage_ns)},
               "host": host_info()}
        if args.cpu_parallel:
            # frame-parallel: one frame per thread (the oracle is reentrant and
            # ctypes releases the GIL for the call)
            from concurrent.futures import ThreadPoolExecutor
            threads = min(len(frames), cpu["host"]["usable_cores"], 16)
            t = time.perf_counter()
            with ThreadPoolExecutor(threads) as ex:
                list(ex.map(oracle.chain, frames))
            par_s = time.perf_counter() - t
            cpu["frame_parallel"] = {"value": round(total_px / par_s / 1e6, 4), "cores": threads,
                                     "sample": f"all {len(frames)} frames of the step, one frame per "
                                               f"thread, {par_s:.1f} s"}

    dropin = None if args.no_dropin else dropin_rate(frames[0], plain[0])

    value = total_px / s_batch / 1e6
    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "Mpixel/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(s_batch * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "+".join(sorted({"u16" if f.desc.bit_depth > 8 else "u8" for f in frames}, reverse=True)),
        "data": "synthetic",
        "config": {
            "workload": describe(frames) + "; " + stages_on(frames),
            "frames_per_step": len(frames),
            "path": "batched persistent kernels, hipGraph replay",
            "launch_shapes": eng.tuning_source or "built-in defaults",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(args.cfg, len(frames), dom),
        },
        "cpu_baseline": cpu,
        "plain_mpx_s": round(total_px / s_plain / 1e6, 1),
        "batched_nograph_mpx_s": round(total_px / s_batch_nograph / 1e6, 1),
        "batched_vs_plain": round(s_plain / s_batch, 3),
        "latency_ms": {"p50": round(p50, 3), "p99": round(p99, 3),
                       "what": f"one {d0.upscaled_width}x{d0.frame_height} frame alone, plain path, "
                               f"enqueue to completion"},
        "stage_ms_per_batch": {k: round(v, 4) for k, v in stage_ms.items()},
        "stage_gbs": {k: round(sb[k] / (stage_ms[k] * 1e-3) / 1e9, 1) if stage_ms[k] else 0
                      for k in stage_ms},
        "chain_algorithmic_gbs": round(chain_gbs, 1),
        "chain_roofline_frac": round(chain_gbs / HBM_PEAK_GBS, 4),
        "byte_exact": {"batched_eq_plain_frames": exact_paths, "frames": len(frames),
                       "oracle_checked": len(pick) if cpu else 0,
                       "oracle_exact": oracle_exact},
        "dropin_path": dropin,
        "device": {"cus": num_cus, "hbm_gb": round(hbm_total / 2**30, 1)},
        "setup_s": round(gen_s, 2),
    }
    print(json.dumps(line))
    eng.close()


if __name__ == "__main__":
    main()
