"""Serving loop: concurrent decode sessions through the library's server
(oxav1_pf_server_*: HBM arena admission, a FIFO of ready frames, lanes that
take what is ready when they free up).

A DecodeSession stands in for one decoder: the CPU produces its pre-filter
frames (seeded mode info and parameters, a pool of distinct frames reused
in order) and its frames arrive as a Poisson process.  With host=True the
frames live in page-locked host memory, as the decoder's CPU stages leave
them: every submission copies its frame to one of the session's input
buffers on the server's upload stream while other frames are on the device,
and the results are copied back to host memory before the frame is done
(frame_driver.rs:790-824, 1098-1116).  Otherwise the pool sits in the
session's input region of the server's arena, uploaded once.  Frame k of a session uses pool
frame k % pool and writes reference slot k % slots; the slots stay resident
in HBM for the session's lifetime (the pre-grain reference store of
frame_driver.rs:1044-1052).

run_serving() drives the loop in real time on the server's clock: it admits
waiting sessions as memory frees up, submits frames as they arrive, polls
the server, and when a session has received all its frames it compares the
frames still held in its slots with the same frames filtered alone, then
closes the session.
"""
import ctypes as C
import heapq
import random

import numpy as np

from . import abi
from .frames import Frame

_ALIGN = 256


def _al(n):
    return (n + _ALIGN - 1) // _ALIGN * _ALIGN


def input_layout(frame):
    """Byte offsets of a frame's pre-filter inputs inside one region:
    [(name, numpy array, offset)], total bytes."""
    d = frame.desc
    st = np.uint8 if d.bit_depth == 8 else np.uint16
    items = [(f"src{p}", np.ascontiguousarray(frame.planes[p].astype(st))) for p in range(d.num_planes)]
    items += [("lf_mi", frame.lf_mi), ("delta_lf", frame.delta_lf), ("cdef_idx", frame.cdef_idx)]
    items += [(f"tx{p}", frame.lf_tx_size[p]) for p in range(3)]
    for p in range(3):
        u = frame.lr_units[p]
        if u is not None:
            items.append((f"lr{p}", np.frombuffer(bytes(u), dtype=np.uint8)))
    out, off = [], 0
    for name, a in items:
        a = np.ascontiguousarray(a)
        out.append((name, a, off))
        off += _al(a.nbytes)
    return out, off


def upload_inputs(frame, base):
    """Copy a frame's inputs to device memory at base; returns its FrameDev
    (outputs left empty: the server points them at a session slot)."""
    lib = abi.load()
    layout, _ = input_layout(frame)
    ptr = {}
    for name, a, off in layout:
        ptr[name] = base + off
        if a.nbytes:
            abi.check(lib.oxav1_pf_dev_upload(base + off, a.ctypes.data, a.nbytes), "dev_upload")
    d = frame.desc
    v = abi.FrameDev()
    for p in range(d.num_planes):
        v.src[p] = ptr[f"src{p}"]
        v.src_pitch[p] = frame.planes[p].shape[1]
    v.lf_mi = ptr["lf_mi"]
    v.delta_lf = ptr["delta_lf"]
    v.cdef_idx = ptr["cdef_idx"]
    for p in range(3):
        v.lf_tx_size[p] = ptr[f"tx{p}"]
        v.lr_units[p] = ptr.get(f"lr{p}")
    return v


class HostArray:
    """A numpy array over page-locked host memory (oxav1_pf_host_alloc), so
    the server's copies from and to it are asynchronous."""

    def __init__(self, shape, dtype):
        self.lib = abi.load()
        dt = np.dtype(dtype)
        n = int(np.prod(shape)) * dt.itemsize
        p = C.c_void_p()
        abi.check(self.lib.oxav1_pf_host_alloc(C.byref(p), max(n, 1)), "host_alloc")
        self.ptr = p.value
        self.a = np.frombuffer((C.c_char * max(n, 1)).from_address(self.ptr), dtype=dt,
                               count=int(np.prod(shape))).reshape(shape)

    def free(self):
        if self.ptr:
            self.a = None
            self.lib.oxav1_pf_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class HostFrame:
    """One pre-filter frame as the CPU stages of a decoder leave it: planes
    (u8/u16) and side information in page-locked host memory."""

    def __init__(self, frame):
        d = frame.desc
        st = np.uint8 if d.bit_depth == 8 else np.uint16
        self.bufs = []

        def put(a, dtype):
            h = HostArray(a.shape, dtype)
            h.a[...] = a
            self.bufs.append(h)
            return h.ptr

        v = abi.FrameHost()
        for p in range(d.num_planes):
            v.src[p] = put(frame.planes[p].astype(st), st)
            v.src_pitch[p] = frame.planes[p].shape[1]
        v.lf_mi = put(frame.lf_mi, np.uint32)
        v.delta_lf = put(frame.delta_lf, np.int8)
        for p in range(3):
            v.lf_tx_size[p] = put(frame.lf_tx_size[p], np.uint8)
            u = frame.lr_units[p]
            if u is not None:
                v.lr_units[p] = put(np.frombuffer(bytes(u), dtype=np.uint8), np.uint8)
        v.cdef_idx = put(frame.cdef_idx, np.int8)
        self.v = v


class Server:
    """ctypes face of oxav1_pf_server_*."""

    def __init__(self, lanes=4, max_batch=16, arena_bytes=0, largest=None):
        self.lib = abi.load()
        abi.check(self.lib.oxav1_pf_init(), "init")
        cfg = abi.ServerConfig()
        cfg.lanes, cfg.max_batch, cfg.arena_bytes = lanes, max_batch, arena_bytes
        self._largest = abi.FrameDesc.from_buffer_copy(largest) if largest is not None else None
        cfg.largest = C.addressof(self._largest) if self._largest is not None else None
        h = C.c_void_p()
        abi.check(self.lib.oxav1_pf_server_create(C.byref(h), C.byref(cfg)), "server_create")
        self.h = h
        self._done = (abi.ServerDone * 1024)()

    def close(self):
        if self.h:
            self.lib.oxav1_pf_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def now_us(self):
        return self.lib.oxav1_pf_server_now_us()

    def open_session(self, shape, ref_slots, input_bytes):
        """(session id, input region pointer, footprint) or None when refused."""
        sid, inp, fp = C.c_uint32(), C.c_void_p(), C.c_uint64()
        st = self.lib.oxav1_pf_server_open_session(self.h, C.byref(shape), ref_slots, input_bytes,
                                                   C.byref(sid), C.byref(inp), C.byref(fp))
        if st == abi.ERR_ADMISSION_REFUSED:
            return None
        abi.check(st, "server_open_session")
        return sid.value, inp.value, fp.value

    def close_session(self, sid):
        st = self.lib.oxav1_pf_server_close_session(self.h, sid)
        if st == abi.ERR_BUSY:
            return False
        abi.check(st, "server_close_session")
        return True

    def submit(self, sid, desc, dev, slot, arrival_us):
        """Ticket, or None when the slot is still busy."""
        t = C.c_uint64()
        st = self.lib.oxav1_pf_server_submit(self.h, sid, C.byref(desc), C.byref(dev), slot, arrival_us,
                                             C.byref(t))
        if st == abi.ERR_BUSY:
            return None
        abi.check(st, "server_submit")
        return t.value

    def submit_host(self, sid, desc, host, slot, arrival_us):
        """Ticket, or None when the slot or every input buffer is busy."""
        t = C.c_uint64()
        st = self.lib.oxav1_pf_server_submit_host(self.h, sid, C.byref(desc), C.byref(host), slot, arrival_us,
                                                  C.byref(t))
        if st == abi.ERR_BUSY:
            return None
        abi.check(st, "server_submit_host")
        return t.value

    def frame_input_bytes(self, desc):
        return int(self.lib.oxav1_pf_server_frame_input_bytes(C.byref(desc)))

    def submit_existing(self, sid, slot, fg, arrival_us):
        """show_existing_frame output of a slot with grain parameters fg;
        ticket, or None when the slot is still busy."""
        t = C.c_uint64()
        st = self.lib.oxav1_pf_server_submit_existing(self.h, sid, slot, C.byref(fg), arrival_us, C.byref(t))
        if st == abi.ERR_BUSY:
            return None
        abi.check(st, "server_submit_existing")
        return t.value

    def poll(self):
        """Finished frames as a list of ServerDone copies."""
        out = []
        while True:
            n = C.c_uint32()
            abi.check(self.lib.oxav1_pf_server_poll(self.h, self._done, len(self._done), C.byref(n)),
                      "server_poll")
            out += [abi.ServerDone.from_buffer_copy(self._done[i]) for i in range(n.value)]
            if n.value < len(self._done):
                return out

    def slot(self, sid, slot, plane, which, shape, dtype):
        """Download one plane of a session slot (which 0 = pre-grain, 1 = grain output)."""
        p, pitch = C.c_void_p(), C.c_uint32()
        abi.check(self.lib.oxav1_pf_server_slot(self.h, sid, slot, plane, which, C.byref(p), C.byref(pitch)),
                  "server_slot")
        assert pitch.value == shape[1]
        a = np.empty(shape, dtype=dtype)
        abi.check(self.lib.oxav1_pf_dev_download(a.ctypes.data, p.value, a.nbytes), "dev_download")
        return a

    def stats(self):
        s = abi.ServerStats()
        abi.check(self.lib.oxav1_pf_server_stats(self.h, C.byref(s)), "server_stats")
        return {k: getattr(s, k) for k, _ in abi.ServerStats._fields_}


class DecodeSession:
    """One synthetic decode stream: shape and parameters drawn from the seed
    (serving mix: 1280x720 / 1920x1080 / 3840x2160, 8 or 10-bit 4:2:0 unless
    a GenConfig is given), a pool of distinct pre-filter frames, Poisson
    arrivals at fps."""

    def __init__(self, seed, n_frames, fps=30.0, pool=8, slots=8, gen_config=None, host=False, in_flight=3):
        """host: the frames stay in page-locked host memory and every
        submission copies one to the device (oxav1_pf_server_submit_host;
        in_flight input buffers per session), and every result is copied
        back to host memory; else the pool is uploaded once at admission."""
        lib = abi.load()
        if gen_config is None:
            gen_config = abi.GenConfig()
            abi.check(lib.oxav1_gen_default_config(4, seed, C.byref(gen_config)), "gen_default_config")
        self.seed = seed
        self.frames = [Frame.generate(gen_config, int(lib.oxav1_gen_splitmix64(seed ^ (k + 1))))
                       for k in range(pool)]
        self.n_frames = n_frames
        self.slots = slots
        self.desc = self.frames[0].desc
        self.pixels = int(self.desc.upscaled_width) * int(self.desc.frame_height)
        self.host = host
        if host:
            self.host_frames = [HostFrame(f) for f in self.frames]
            st = np.uint8 if self.desc.bit_depth == 8 else np.uint16
            grain = bool(self.desc.fg.apply_grain)
            self.host_ref = [[HostArray(self.frames[0].out_shape(p), st) for p in range(self.desc.num_planes)]
                             for _ in range(slots)]
            self.host_out = [[HostArray(self.frames[0].out_shape(p), st) for p in range(self.desc.num_planes)]
                             if grain else None for _ in range(slots)]
            per = int(lib.oxav1_pf_server_frame_input_bytes(C.byref(self.desc)))
            self.input_bytes = per * in_flight
            self.input_offsets = None
        else:
            layouts = [input_layout(f)[1] for f in self.frames]
            self.input_offsets = np.cumsum([0] + layouts[:-1]).tolist()
            self.input_bytes = int(sum(layouts))
        rng = random.Random(seed)
        t, self.offsets_us = 0.0, []
        for _ in range(n_frames):
            t += rng.expovariate(fps) * 1e6
            self.offsets_us.append(t)
        self.sid = None
        self.devs = None
        self.start_us = None
        self.next = 0
        self.done = 0
        self.submitted = 0
        self.latency_ms = []
        self.closed = False

    @property
    def sample_dtype(self):
        return np.uint8 if self.desc.bit_depth == 8 else np.uint16

    def out_shape(self, p):
        return self.frames[0].out_shape(p)

    def admit(self, server):
        """Open the session on the server and put its pre-filter frames in
        its input region; False when the server refuses it (HBM full).
        The session's clock (first arrival) starts at start()."""
        r = server.open_session(self.desc, self.slots, self.input_bytes)
        if r is None:
            return False
        self.sid, base, self.footprint = r
        if not self.host:
            self.devs = [upload_inputs(f, base + off) for f, off in zip(self.frames, self.input_offsets)]
        return True

    def start(self, now_us):
        self.start_us = now_us

    def frame_job(self, k):
        j = k % len(self.frames)
        return self.frames[j].desc, self.devs[j], k % self.slots

    def frame_job_host(self, k):
        """(desc, FrameHost with this frame's inputs and its slot's host
        output buffers, slot)."""
        j, slot = k % len(self.frames), k % self.slots
        v = abi.FrameHost.from_buffer_copy(self.host_frames[j].v)
        for p in range(self.desc.num_planes):
            v.ref_out[p] = self.host_ref[slot][p].ptr
            if self.host_out[slot] is not None:
                v.out[p] = self.host_out[slot][p].ptr
        return self.frames[j].desc, v, slot

    def submit(self, server, k, arrival_us):
        if self.host:
            desc, v, slot = self.frame_job_host(k)
            return server.submit_host(self.sid, desc, v, slot, arrival_us)
        desc, dev, slot = self.frame_job(k)
        return server.submit(self.sid, desc, dev, slot, arrival_us)

    def held_frames(self):
        """(slot, pool index) of the frames the slots hold once every frame is done."""
        last = {}
        for k in range(self.n_frames):
            last[k % self.slots] = k % len(self.frames)
        return sorted(last.items())


def run_alone(frame, engine=None):
    """The frame filtered by itself: its own launches, one per stage (the
    plain path), on `engine` or a fresh one.  Returns (ref_out planes,
    grain output planes)."""
    from .engine import DeviceFrame, Engine
    df = DeviceFrame(frame)
    eng = engine or Engine(max_frames=1)
    try:
        eng.prepare([df])
        eng.run(Engine.MODE_PLAIN)
        eng.sync()
        ref, out = df.result()
    finally:
        if engine is None:
            eng.close()
        df.free()
    return ref, out


def expected_outputs(sessions, engine=None):
    """{(session index, pool index): run_alone(...)} for the frames each
    session's slots hold at its end."""
    from .engine import Engine
    eng = engine or Engine(max_frames=1)
    try:
        return {(i, j): run_alone(s.frames[j], eng)
                for i, s in enumerate(sessions) for _, j in s.held_frames()}
    finally:
        if engine is None:
            eng.close()


def run_serving(server, sessions, expected=None, max_seconds=120.0):
    """Drive the sessions through the server until every frame is done.

    Sessions are admitted in order as the arena has room; a session's
    frames start arriving once its inputs are uploaded.  expected:
    optional {(session index, pool index): (ref planes, out planes)} to
    compare each session's held slots with once it is done.  A finished
    session is compared and closed at once only when another session waits
    for its memory; otherwise after the loop, so that slot downloads do not
    stall the frames of the sessions still running.
    Returns a dict of results (latencies, throughput, mismatches, stats)."""
    waiting = list(range(len(sessions)))
    events = []  # (arrival_us, session index, frame index)
    lat_ms = []
    dev_ms = []
    queue_ms = []
    batches = []
    mismatches = []
    verified = 0
    finished = []
    done_px = 0
    t0 = server.now_us()
    first_arrival = None
    last_done = t0
    by_sid = {}
    peak_open = 0
    open_now = 0

    def admit_waiting():
        nonlocal peak_open, open_now
        fresh = []
        while waiting:
            s = sessions[waiting[0]]
            if not s.admit(server):
                break
            fresh.append(waiting.pop(0))
            by_sid[s.sid] = fresh[-1]
            open_now += 1
        now = server.now_us()
        for i in fresh:
            s = sessions[i]
            s.start(now)
            for k, off in enumerate(s.offsets_us):
                heapq.heappush(events, (s.start_us + off, i, k))
        peak_open = max(peak_open, open_now)

    def finish(i):
        nonlocal verified, open_now
        s = sessions[i]
        if expected is not None:
            for slot, j in s.held_frames():
                ref_e, out_e = expected[(i, j)]
                for p in range(s.desc.num_planes):
                    got = server.slot(s.sid, slot, p, 0, s.out_shape(p), s.sample_dtype)
                    ok = np.array_equal(got.astype(np.int32), ref_e[p])
                    if ok and s.desc.fg.apply_grain:
                        got = server.slot(s.sid, slot, p, 1, s.out_shape(p), s.sample_dtype)
                        ok = np.array_equal(got.astype(np.int32), out_e[p])
                    if ok and s.host:  # the copies handed back to the host
                        ok = np.array_equal(s.host_ref[slot][p].a.astype(np.int32), ref_e[p])
                        if ok and s.host_out[slot] is not None:
                            ok = np.array_equal(s.host_out[slot][p].a.astype(np.int32), out_e[p])
                    if not ok:
                        mismatches.append((i, slot, j, p))
                verified += 1
        assert server.close_session(s.sid)
        open_now -= 1

    admit_waiting()
    retry = []
    n_done = 0
    while len(finished) < len(sessions):
        now = server.now_us()
        if (now - t0) * 1e-6 > max_seconds:
            raise TimeoutError(f"serving loop exceeded {max_seconds} s "
                               f"({len(finished)} of {len(sessions)} sessions done)")
        # frames whose slot was busy at arrival go first, in arrival order
        pend = retry
        retry = []
        while events and events[0][0] <= now:
            pend.append(heapq.heappop(events))
        for arr, i, k in pend:
            s = sessions[i]
            if s.submit(server, k, arr) is None:
                retry.append((arr, i, k))
                continue
            if first_arrival is None:
                first_arrival = arr
        for r in server.poll():
            i = by_sid[r.session]
            s = sessions[i]
            if r.status != abi.OK:
                raise abi.Oxav1Error(r.status, f"session {i} frame in lane batch of {r.batch}")
            s.done += 1
            n_done += 1
            s.latency_ms.append((r.done_us - r.arrival_us) * 1e-3)
            lat_ms.append((r.done_us - r.arrival_us) * 1e-3)
            dev_ms.append((r.done_us - r.launch_us) * 1e-3)
            queue_ms.append((r.launch_us - r.arrival_us) * 1e-3)
            batches.append(r.batch)
            done_px += s.pixels
            last_done = max(last_done, r.done_us)
            if s.done == s.n_frames:
                finished.append(i)
                if waiting:
                    finish(i)
                    s.closed = True
                    admit_waiting()
    span_s = max(last_done - (first_arrival if first_arrival is not None else t0), 1.0) * 1e-6
    for i in finished:
        if not sessions[i].closed:
            finish(i)
            sessions[i].closed = True
    lat = np.sort(np.array(lat_ms)) if lat_ms else np.zeros(1)
    return {
        "frames": len(lat_ms),
        "pixels": done_px,
        "span_s": span_s,
        "mpx_s": done_px / span_s / 1e6,
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "max_ms": float(lat[-1]),
        "queue_p99_ms": float(np.percentile(np.array(queue_ms), 99)) if queue_ms else 0.0,
        "launch_to_done_p50_ms": float(np.percentile(np.array(dev_ms), 50)) if dev_ms else 0.0,
        "launch_to_done_p99_ms": float(np.percentile(np.array(dev_ms), 99)) if dev_ms else 0.0,
        "mean_batch": float(np.mean(batches)) if batches else 0.0,
        "max_batch": int(max(batches)) if batches else 0,
        "verified_frames": verified,
        "mismatches": mismatches,
        "peak_sessions_open": peak_open,
        "stats": server.stats(),
    }
