"""Python-level calls of the five drop-in entry points.

`FilterCalls` wraps any library exporting the five functions with the
include/oxav1_pf.h signatures (the HIP library here; the CPU oracle binds
the same class in oracle/oracle.py), taking and returning numpy int32
planes so results of different implementations compare directly.
"""
import ctypes as C

import numpy as np

from . import abi
from .frames import planes_struct

GPU_NAMES = {
    "lf": "oxav1_pf_loop_filter_frame",
    "cdef": "oxav1_pf_cdef_frame",
    "sr": "oxav1_pf_upscale_frame",
    "lr": "oxav1_pf_loop_restoration_frame",
    "fg": "oxav1_pf_film_grain_synthesis",
}


class FilterCalls:
    def __init__(self, lib, names, status_returning=True):
        self.lib = lib
        self.fn = {k: getattr(lib, v) for k, v in names.items()}
        self.status_returning = status_returning

    def _st(self, ret, what):
        if self.status_returning:
            abi.check(ret, what)
        return ret

    # loop_filter_frame (loop_filter.rs:271), in place -> returns new arrays
    # (inplace=True filters the caller's int32 planes themselves)
    def loop_filter_frame(self, frame, planes, ctx=None, inplace=False):
        if inplace:
            assert all(p.dtype == np.int32 and p.flags.c_contiguous for p in planes)
            out = list(planes)
        else:
            out = [np.ascontiguousarray(p, dtype=np.int32).copy() for p in planes]
        c, g, keep = frame.lf_args()
        if ctx is not None:
            c = ctx
        ps = planes_struct(out)
        self._st(self.fn["lf"](C.byref(c), C.byref(g), ps, len(out)), "loop_filter_frame")
        return out

    # cdef_frame (cdef.rs:180): dst starts as `dst_init` (default: copy of
    # src), in `out` when given (preallocated int32 planes of src's shapes)
    def cdef_frame(self, frame, src, dst_init=None, out=None):
        s = [np.ascontiguousarray(p, dtype=np.int32) for p in src]
        init = dst_init if dst_init is not None else s
        if out is not None:
            for o, p in zip(out, init):
                np.copyto(o, p)
            d = list(out)
        else:
            d = [p.copy() for p in init]
        c, g, keep = frame.cdef_args()
        ss, ds = planes_struct(s), planes_struct(d)
        self._st(self.fn["cdef"](C.byref(c), C.byref(g), ss, len(s), ds, len(d)), "cdef_frame")
        return d

    # upscale_frame (superres.rs:282): returns (status, outputs)
    def upscale_frame(self, ctx, inp, out_shapes, out=None):
        i = [np.ascontiguousarray(p, dtype=np.int32) for p in inp]
        o = list(out) if out is not None else [np.zeros(s, dtype=np.int32) for s in out_shapes]
        st = self.fn["sr"](C.byref(ctx), planes_struct(i), len(i), planes_struct(o), len(o))
        return st, o

    # loop_restoration_frame (loop_restoration.rs:268)
    # (the call itself starts LrFrame as the CDEF planes: `out`, when given,
    # is used as it is)
    def loop_restoration_frame(self, frame, curr, cdef, ctx=None, grids=None, lr_init=None, out=None):
        cu = [np.ascontiguousarray(p, dtype=np.int32) for p in curr]
        cd = [np.ascontiguousarray(p, dtype=np.int32) for p in cdef]
        if out is not None:
            lr = list(out)
        else:
            lr = [p.copy() for p in lr_init] if lr_init is not None else [np.zeros_like(p) for p in cd]
        if frame is not None:
            c, g, keep = frame.lr_args()
        if ctx is not None:
            c = ctx
        if grids is not None:
            g = grids
        self._st(self.fn["lr"](C.byref(c), C.byref(g), planes_struct(cu), len(cu),
                               planes_struct(cd), len(cd), planes_struct(lr), len(lr)),
                 "loop_restoration_frame")
        return lr

    # film_grain_synthesis (film_grain.rs:848), in place -> returns new arrays
    def film_grain_synthesis(self, fg, bit_depth, num_planes, sx, sy, mc, planes, out=None):
        if out is not None:
            for o, p in zip(out, planes):
                np.copyto(o, p)
            out = list(out)
        else:
            out = [np.ascontiguousarray(p, dtype=np.int32).copy() for p in planes]
        self._st(self.fn["fg"](C.byref(fg), bit_depth, num_planes, sx, sy, mc,
                               planes_struct(out), len(out)), "film_grain_synthesis")
        return out


_gpu = None


def gpu():
    """FilterCalls over the HIP library (raises if it is not built)."""
    global _gpu
    if _gpu is None:
        _gpu = FilterCalls(abi.load(), GPU_NAMES, status_returning=True)
    return _gpu


def driver_buffers(frame):
    """int32 planes the frame driver reuses from frame to frame, as a decoder
    reuses its frame buffers (allocated and touched once)."""
    d = frame.desc
    ins = [p.shape for p in frame.planes]
    outs = [frame.out_shape(p) for p in range(d.num_planes)]

    def mk(shapes):
        return [np.ones(sh, dtype=np.int32) for sh in shapes]
    return {"rec": mk(ins), "cdef": mk(ins), "upc": mk(outs), "upd": mk(outs), "lr": mk(outs), "out": mk(outs)}


def frame_driver_chain(calls, frame, stage_s=None, bufs=None):
    """The post-filter part of the reference's frame driver
    (frame_driver.rs:790-1116) expressed with the five drop-in calls, the
    way the decoder would drive them once bound to the C ABI: deblock in
    place; keep a deblocked copy when loop restoration is on; CDEF from a
    copy into the live planes; superres of the CDEF set and of the
    deblocked set; loop restoration reading the deblocked set and a copy of
    the CDEF set into the live planes; crop; film grain on a separate
    output.  `calls` is a FilterCalls (the HIP library or the oracle).
    Returns (pre-grain reference planes, output planes); stage_s, when a
    list, receives the wall time of each of the five stages.  bufs (from
    driver_buffers) makes the stages write into reused planes: the frame's
    reconstruction is copied into bufs["rec"] before the first stage's
    clock starts (a decoder filters its reconstruction in place), and the
    copies the driver itself makes (the deblocked planes kept for loop
    restoration, the pre-grain reference kept beside the grain output) go
    into reused planes."""
    import time
    d = frame.desc
    np_ = d.num_planes
    times = []
    b = bufs or {}
    if bufs is not None:
        for o, p in zip(bufs["rec"], frame.planes):
            np.copyto(o, p)
        rec = list(bufs["rec"])
    t = time.perf_counter()
    if bufs is None:
        rec = [p.copy() for p in frame.planes]
    if d.apply_lf:
        rec = calls.loop_filter_frame(frame, rec, inplace=True)
    times.append(time.perf_counter() - t)
    t = time.perf_counter()
    uses_lr = bool(d.lr.uses_lr)
    deb = rec
    cur = calls.cdef_frame(frame, rec, out=b.get("cdef")) if d.apply_cdef else rec
    times.append(time.perf_counter() - t)
    t = time.perf_counter()
    shapes = [frame.out_shape(p) for p in range(np_)]
    if d.use_superres and d.upscaled_width > d.frame_width:
        ctx = frame.sr_ctx()
        st, upc = calls.upscale_frame(ctx, cur, shapes, out=b.get("upc"))
        abi.check(st, "upscale_frame (CDEF set)")
        upd = upc
        if uses_lr:
            st, upd = calls.upscale_frame(ctx, deb, shapes, out=b.get("upd"))
            abi.check(st, "upscale_frame (deblocked set)")
    else:
        upc, upd = cur, deb
    times.append(time.perf_counter() - t)
    t = time.perf_counter()
    if uses_lr and bufs is not None:
        fin = calls.loop_restoration_frame(frame, upd, upc, out=b["lr"])
    else:
        fin = calls.loop_restoration_frame(frame, upd, upc, lr_init=upc) if uses_lr else upc
    times.append(time.perf_counter() - t)
    t = time.perf_counter()
    ref = [np.ascontiguousarray(fin[p][:shapes[p][0], :shapes[p][1]]) for p in range(np_)]
    out = ref
    if d.fg.apply_grain:
        out = calls.film_grain_synthesis(d.fg, d.bit_depth, np_, d.subsampling_x, d.subsampling_y,
                                         d.matrix_coefficients, ref, out=b.get("out"))
    times.append(time.perf_counter() - t)
    if stage_s is not None:
        stage_s[:] = times
    return ref, out
