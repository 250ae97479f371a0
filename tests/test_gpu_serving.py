"""GPU serving loop (oxav1_pf_server_*): decode sessions admitted by HBM
footprint, frames batched across sessions onto lanes as they arrive, and
every frame's bytes the same in the mix as filtered alone (and equal to the
oracle's chain, frame_driver.rs:790-1116)."""
import numpy as np
import pytest

from conftest import gen_config

pytestmark = pytest.mark.gpu

FG = dict(enable_fg=1, fg_num_y_points=6, fg_ar_lag=2, fg_chroma_points=3, fg_overlap=1)
MIX = [
    dict(),
    dict(bit_depth=10, noise_amp=32, upscaled_width=160, frame_height=96, **FG),
    dict(bit_depth=10, noise_amp=32, superres_denom=12, upscaled_width=192, frame_height=72),
    dict(subsampling_x=0, subsampling_y=0, upscaled_width=120, frame_height=64, **FG),
    dict(bit_depth=12, noise_amp=128, subsampling_y=0, upscaled_width=136, frame_height=88),
    dict(mono=1, upscaled_width=104, frame_height=56),
]


def _sessions(cfgs, n_frames, fps, seed0=0x5E55, host=False):
    from oxav1.serving import DecodeSession
    return [DecodeSession(seed0 + i, n_frames=n_frames, fps=fps, pool=n_frames, slots=n_frames,
                          gen_config=gen_config(**kw), host=host) for i, kw in enumerate(cfgs)]


def _assert_alone_is_oracle(sessions, expected, oracle_mod):
    for (i, j), (ref, out) in expected.items():
        ref_o, out_o, _ = oracle_mod.chain(sessions[i].frames[j])
        for p in range(sessions[i].desc.num_planes):
            assert np.array_equal(ref[p], ref_o[p]), f"session {i} frame {j} plane {p}: alone != oracle"
            assert np.array_equal(out[p], out_o[p]), f"session {i} frame {j} plane {p}: alone grain != oracle"


def test_serving_mix_equals_alone(oracle_mod):
    """Six sessions of different shapes, depths and subsamplings, bursty
    arrivals (lanes busy, so batches form): every frame is kept in its own
    slot and compared with the same frame filtered alone."""
    from oxav1.serving import Server, expected_outputs, run_serving
    sessions = _sessions(MIX, n_frames=4, fps=3000.0)
    expected = expected_outputs(sessions)
    _assert_alone_is_oracle(sessions, expected, oracle_mod)
    srv = Server(lanes=3, max_batch=6)
    try:
        res = run_serving(srv, sessions, expected, max_seconds=60)
    finally:
        srv.close()
    print("serving mix:", {k: v for k, v in res.items() if k != "stats"}, res["stats"])
    assert res["frames"] == 4 * len(MIX)
    assert res["verified_frames"] == 4 * len(MIX)
    assert res["mismatches"] == []
    assert res["stats"]["frames_done"] == res["stats"]["frames_submitted"] == 4 * len(MIX)


def test_serving_host_frames_mix(oracle_mod):
    """The same mix with every frame in host memory: each submission copies
    the frame to the device while other frames are being filtered, and the
    results come back to host memory; bursty arrivals with two input
    buffers per session, so submissions also wait for free buffers.  Every
    host copy and every slot equals the frame filtered alone."""
    from oxav1.serving import DecodeSession, Server, expected_outputs, run_serving
    sessions = [DecodeSession(0x5E57 + i, n_frames=6, fps=3000.0, pool=3, slots=3,
                              gen_config=gen_config(**kw), host=True, in_flight=2)
                for i, kw in enumerate(MIX)]
    expected = expected_outputs(sessions)
    _assert_alone_is_oracle(sessions, expected, oracle_mod)
    srv = Server(lanes=3, max_batch=6)
    try:
        res = run_serving(srv, sessions, expected, max_seconds=60)
    finally:
        srv.close()
    print("serving host frames:", {k: v for k, v in res.items() if k != "stats"}, flush=True)
    assert res["frames"] == 6 * len(MIX)
    assert res["verified_frames"] == 3 * len(MIX)
    assert res["mismatches"] == []


def test_serving_admission_by_footprint():
    """An arena that holds two sessions: the others wait and are admitted
    as sessions finish and free their slots; results still match alone."""
    from oxav1.serving import Server, expected_outputs, run_serving
    sessions = _sessions([MIX[1]] * 5, n_frames=3, fps=2000.0, seed0=0xAD17)
    expected = expected_outputs(sessions)
    probe = Server(lanes=1, max_batch=1, arena_bytes=1 << 26)
    try:
        sid, _, fp = probe.open_session(sessions[0].desc, sessions[0].slots, sessions[0].input_bytes)
        assert probe.close_session(sid)
    finally:
        probe.close()
    srv = Server(lanes=2, max_batch=4, arena_bytes=int(fp * 2.5))
    try:
        a = srv.open_session(sessions[0].desc, 3, sessions[0].input_bytes)
        b = srv.open_session(sessions[0].desc, 3, sessions[0].input_bytes)
        assert a is not None and b is not None
        assert srv.open_session(sessions[0].desc, 3, sessions[0].input_bytes) is None  # refused
        assert srv.close_session(a[0]) and srv.close_session(b[0])
        assert srv.stats()["arena_used"] == 0
        res = run_serving(srv, sessions, expected, max_seconds=60)
    finally:
        srv.close()
    assert res["peak_sessions_open"] == 2
    assert res["frames"] == 15 and res["verified_frames"] == 15
    assert res["mismatches"] == []


def test_serving_slot_busy_until_done():
    """A slot with a frame queued or in flight refuses another frame, and
    its session cannot close until the frame is done."""
    from oxav1.serving import Server
    s = _sessions([MIX[0]], n_frames=1, fps=30.0)[0]
    srv = Server(lanes=1, max_batch=1)
    try:
        assert s.admit(srv)
        desc, dev, slot = s.frame_job(0)
        assert srv.submit(s.sid, desc, dev, slot, srv.now_us()) is not None
        assert srv.submit(s.sid, desc, dev, slot, srv.now_us()) is None
        assert not srv.close_session(s.sid)
        done = []
        t0 = srv.now_us()
        while not done:
            done = srv.poll()
            assert srv.now_us() - t0 < 30e6, "frame did not finish"
        assert len(done) == 1 and done[0].status == 0 and done[0].batch == 1
        assert srv.submit(s.sid, desc, dev, slot, srv.now_us()) is not None
        while srv.stats()["frames_done"] < 2:
            srv.poll()
            assert srv.now_us() - t0 < 30e6, "frame did not finish"
        assert srv.close_session(s.sid)
    finally:
        srv.close()


def _drive(srv, n_expected, timeout_us=30e6):
    done, t0 = [], srv.now_us()
    while len(done) < n_expected:
        done += srv.poll()
        assert srv.now_us() - t0 < timeout_us, "frames did not finish"
    return done


def test_serving_show_existing_frame(oracle_mod):
    """show_existing_frame output (frame_driver.rs:1489-1537): grain from a
    slot's saved parameters applied to the stored pre-grain frame, into the
    slot's grain buffer; equals the oracle's film_grain_synthesis of the
    stored frame, and the stored frame stays as it was.  Also the no-grain
    case (a copy)."""
    from oxav1 import abi
    from oxav1.serving import Server, expected_outputs
    sessions = _sessions([MIX[1], MIX[3]], n_frames=2, fps=30.0, seed0=0x5E0E)
    expected = expected_outputs(sessions)
    calls = oracle_mod.calls()
    srv = Server(lanes=2, max_batch=4)
    try:
        for s in sessions:
            assert s.admit(srv)
            for k in range(2):
                desc, dev, slot = s.frame_job(k)
                assert srv.submit(s.sid, desc, dev, slot, srv.now_us()) is not None
        assert all(r.status == 0 for r in _drive(srv, 4))
        for i, s in enumerate(sessions):
            d = s.desc
            fg = abi.FilmGrainParams.from_buffer_copy(d.fg)
            fg.grain_seed = (fg.grain_seed * 7 + 1234) & 0xFFFF  # the slot's saved parameters, reseeded
            for slot in range(2):
                assert srv.submit_existing(s.sid, slot, fg, srv.now_us()) is not None
            assert srv.submit_existing(s.sid, 0, fg, srv.now_us()) is None  # busy until done
            assert all(r.status == 0 for r in _drive(srv, 2))
            for slot in range(2):
                ref_e, _ = expected[(i, slot)]
                want = calls.film_grain_synthesis(fg, d.bit_depth, d.num_planes, d.subsampling_x,
                                                  d.subsampling_y, d.matrix_coefficients, ref_e)
                for p in range(d.num_planes):
                    stored = srv.slot(s.sid, slot, p, 0, s.out_shape(p), s.sample_dtype).astype(np.int32)
                    shown = srv.slot(s.sid, slot, p, 1, s.out_shape(p), s.sample_dtype).astype(np.int32)
                    assert np.array_equal(stored, ref_e[p]), (i, slot, p, "stored frame changed")
                    assert np.array_equal(shown, want[p]), (i, slot, p, "shown frame != oracle grain")
            off = abi.FilmGrainParams.from_buffer_copy(fg)
            off.apply_grain = 0
            assert srv.submit_existing(s.sid, 1, off, srv.now_us()) is not None
            _drive(srv, 1)
            for p in range(d.num_planes):
                shown = srv.slot(s.sid, 1, p, 1, s.out_shape(p), s.sample_dtype).astype(np.int32)
                assert np.array_equal(shown, expected[(i, 1)][0][p]), (i, p, "no-grain output != stored")
        for s in sessions:
            assert srv.close_session(s.sid)
    finally:
        srv.close()
