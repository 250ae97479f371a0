"""GPU parity at the benchmark configurations' real sizes:

- the mixed batch (64 frames: 16 each at 640x360, 1280x720, 1920x1080 and
  3840x2160, 8/10-bit, 4:2:0/4:2:2/4:4:4 drawn per frame): batched with
  hipGraph replays equals the plain one-launch-per-frame path on all 64
  frames, and the oracle on one frame of every (resolution, subsampling)
  pair present;
- the serving mix (32 sessions drawn at 720p/1080p/4K, 8/10-bit 4:2:0,
  two frames each, Poisson arrivals): every frame in the mix equals the
  same frame filtered alone, and alone equals the oracle on one frame of
  every (resolution, bit depth) pair present;
- the drop-in calls' chunked staging: a frame whose luma plane spans three
  4 MiB staging chunks (not a multiple of the chunk) through the frame
  driver's sequence of the five calls equals the oracle's chain.
"""
import hashlib
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

THREADS = min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else 8)


def _digest(planes):
    h = hashlib.sha256()
    for p in planes:
        h.update(np.ascontiguousarray(p).tobytes())
    return h.hexdigest()


def _oracle_all(oracle_mod, frames):
    with ThreadPoolExecutor(THREADS) as ex:
        return list(ex.map(oracle_mod.chain, frames))


def test_cfg3_mixed_batch_64_frames(oracle_mod):
    from oxav1 import Frame, workload_seeds
    from oxav1.engine import DeviceFrame, Engine
    seeds = workload_seeds(3, 64)
    frames = [Frame.generate(3, s) for s in seeds]
    shapes = {(f.desc.upscaled_width, f.desc.frame_height) for f in frames}
    assert shapes == {(640, 360), (1280, 720), (1920, 1080), (3840, 2160)}
    dfs = [DeviceFrame(f) for f in frames]
    eng = Engine(max_frames=len(dfs))
    try:
        eng.prepare(dfs)
        eng.run(2)
        eng.sync()
        batched = [_digest(sum(df.result(), [])) for df in dfs]
        for _ in range(2):  # graph replays with cleared outputs reproduce every byte
            for df in dfs:
                df.clear_outputs()
            eng.run(2)
            eng.sync()
            assert [_digest(sum(df.result(), [])) for df in dfs] == batched
        for df in dfs:
            df.clear_outputs()
        eng.run(0)
        eng.sync()
        plain = [_digest(sum(df.result(), [])) for df in dfs]
        bad = [i for i, (a, b) in enumerate(zip(batched, plain)) if a != b]
        assert not bad, f"batched != plain on frames {bad}"
        # the oracle on one frame per (resolution, subsampling) present
        pick, seen = [], set()
        for i, f in enumerate(frames):
            key = (f.desc.upscaled_width, f.desc.subsampling_x, f.desc.subsampling_y)
            if key not in seen:
                seen.add(key)
                pick.append(i)
        assert len(pick) >= 8
        print(f"cfg 3: oracle on {len(pick)} frames: "
              f"{[(frames[i].desc.upscaled_width, frames[i].desc.bit_depth, frames[i].desc.subsampling_x, frames[i].desc.subsampling_y) for i in pick]}",
              flush=True)
        outs = _oracle_all(oracle_mod, [frames[i] for i in pick])
        for i, (ref_o, out_o, _) in zip(pick, outs):
            assert _digest(ref_o + out_o) == plain[i], f"frame {i} != oracle"
    finally:
        eng.close()
        for df in dfs:
            df.free()


@pytest.mark.parametrize("host", [False, True], ids=["device_resident", "host_frames"])
def test_cfg4_serving_32_sessions(host, oracle_mod):
    """host_frames: every frame is copied from page-locked host memory when
    it is submitted and its result copied back (the decoder's CPU hand-off)."""
    from oxav1.serving import DecodeSession, Server, expected_outputs, run_serving
    sessions = [DecodeSession(0x5E4F00 + i, n_frames=2, fps=30.0, pool=2, slots=2, host=host)
                for i in range(32)]
    kinds = {(s.desc.upscaled_width, s.desc.bit_depth) for s in sessions}
    assert {w for w, _ in kinds} == {1280, 1920, 3840}
    expected = expected_outputs(sessions)
    # alone == oracle on one held frame of every (resolution, bit depth) present
    pick, seen = [], set()
    for i, s in enumerate(sessions):
        key = (s.desc.upscaled_width, s.desc.bit_depth)
        if key not in seen:
            seen.add(key)
            pick.append((i, s.held_frames()[0][1]))
    assert len(pick) >= 5
    outs = _oracle_all(oracle_mod, [sessions[i].frames[j] for i, j in pick])
    for (i, j), (ref_o, out_o, _) in zip(pick, outs):
        ref_e, out_e = expected[(i, j)]
        assert _digest(ref_e + out_e) == _digest(ref_o + out_o), f"session {i} frame {j}: alone != oracle"
    srv = Server(lanes=4, max_batch=16)
    try:
        res = run_serving(srv, sessions, expected, max_seconds=120)
    finally:
        srv.close()
    print("cfg 4 serving:", {k: v for k, v in res.items() if k != "stats"}, flush=True)
    assert res["frames"] == 64 and res["verified_frames"] == 64
    assert res["mismatches"] == []


def test_dropin_chunked_staging_equals_oracle(oracle_mod):
    """1928x1090 10-bit 4:2:0: the padded luma plane is 1928x1096 int32 =
    8.45 MB, two full 4 MiB staging chunks and a tail, and every call of
    the frame driver's sequence stages it that way in both directions."""
    from conftest import gen_config
    from oxav1 import Frame, dropin
    f = Frame.generate(gen_config(upscaled_width=1928, frame_height=1090, bit_depth=10,
                                  noise_amp=32, enable_fg=1, fg_num_y_points=8,
                                  fg_ar_lag=2, fg_chroma_points=4, fg_overlap=1), 0xC4A9)
    assert f.planes[0].nbytes > 2 * (4 << 20) and f.planes[0].nbytes % (4 << 20)
    ref_g, out_g = dropin.frame_driver_chain(dropin.gpu(), f)
    ref_o, out_o, _ = oracle_mod.chain(f)
    for p in range(f.num_planes):
        assert np.array_equal(ref_g[p], ref_o[p]), f"plane {p}: pre-grain != oracle"
        assert np.array_equal(out_g[p], out_o[p]), f"plane {p}: grain output != oracle"
