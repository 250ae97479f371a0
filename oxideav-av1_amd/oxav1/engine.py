"""Device-resident chain engine over HBM buffers.

Device memory comes from the library's own allocator (oxav1_pf_dev_*), so
one HIP runtime serves the whole process; every filter runs in the
library's HIP kernels on the engine's stream.
"""
import ctypes as C
import os

import numpy as np

from . import abi


class DevBuffer:
    """One device allocation holding a numpy-shaped array."""

    def __init__(self, shape, dtype, data=None):
        self.lib = abi.load()
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        abi.check(self.lib.oxav1_pf_dev_alloc(C.byref(p), self.nbytes), "dev_alloc")
        self.ptr = p.value
        if data is not None:
            self.upload(data)

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        abi.check(self.lib.oxav1_pf_dev_upload(self.ptr, a.ctypes.data, self.nbytes), "dev_upload")

    def download(self):
        out = np.empty(self.shape, dtype=self.dtype)
        abi.check(self.lib.oxav1_pf_dev_download(out.ctypes.data, self.ptr, self.nbytes), "dev_download")
        return out

    def fill(self, byte_value=0):
        abi.check(self.lib.oxav1_pf_dev_fill(self.ptr, byte_value, self.nbytes), "dev_fill")

    def free(self):
        if self.ptr:
            self.lib.oxav1_pf_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_info():
    lib = abi.load()
    cus, free, total = C.c_int(), C.c_size_t(), C.c_size_t()
    abi.check(lib.oxav1_pf_dev_info(C.byref(cus), C.byref(free), C.byref(total)), "dev_info")
    return cus.value, free.value, total.value


class DeviceFrame:
    """One frame's inputs and outputs in HBM, plus its FrameDesc/FrameDev."""

    def __init__(self, frame, fg_out=True, pad_pitch=False):
        """pad_pitch: the caller's planes (input, pre-grain and grain output)
        get a row pitch of the width rounded up to a multiple of 4 plus 4
        samples, the padding filled with the largest sample value, as a
        decoder's frame buffers with aligned rows would be."""
        self.frame = frame
        d = frame.desc
        self.desc = abi.FrameDesc.from_buffer_copy(d)
        st = np.uint8 if d.bit_depth == 8 else np.uint16
        self.sample_dtype = st
        np_ = d.num_planes
        self.cols = {}

        def plane(shape, data=None):
            rows, cols = shape
            pitch = ((cols + 3) // 4) * 4 + 4 if pad_pitch else cols
            if pitch == cols:
                buf = DevBuffer(shape, st, None if data is None else data.astype(st))
            else:
                full = np.full((rows, pitch), (1 << d.bit_depth) - 1, dtype=st)
                if data is not None:
                    full[:, :cols] = data
                buf = DevBuffer((rows, pitch), st, full)
            self.cols[id(buf)] = cols
            return buf

        self.src = [plane(frame.planes[p].shape, frame.planes[p]) for p in range(np_)]
        self.ref_out = [plane(frame.out_shape(p)) for p in range(np_)]
        self.out = ([plane(frame.out_shape(p)) for p in range(np_)]
                    if (fg_out and d.fg.apply_grain) else None)
        self.lf_mi = DevBuffer(frame.lf_mi.shape, np.uint32, frame.lf_mi)
        self.delta_lf = DevBuffer(frame.delta_lf.shape, np.int8, frame.delta_lf)
        self.tx = [DevBuffer(frame.lf_tx_size[p].shape, np.uint8, frame.lf_tx_size[p]) for p in range(3)]
        self.cdef_idx = DevBuffer(frame.cdef_idx.shape, np.int8, frame.cdef_idx)
        self.units = []
        for p in range(3):
            u = frame.lr_units[p]
            if u is None:
                self.units.append(None)
            else:
                raw = np.frombuffer(bytes(u), dtype=np.uint8)
                self.units.append(DevBuffer(raw.shape, np.uint8, raw))
        v = abi.FrameDev()
        for p in range(np_):
            v.src[p] = self.src[p].ptr
            v.src_pitch[p] = self.src[p].shape[1]
            v.ref_out[p] = self.ref_out[p].ptr
            v.ref_pitch[p] = self.ref_out[p].shape[1]
            if self.out is not None:
                v.out[p] = self.out[p].ptr
                v.out_pitch[p] = self.out[p].shape[1]
        v.lf_mi = self.lf_mi.ptr
        v.delta_lf = self.delta_lf.ptr
        for p in range(3):
            v.lf_tx_size[p] = self.tx[p].ptr
            v.lr_units[p] = self.units[p].ptr if self.units[p] is not None else None
        v.cdef_idx = self.cdef_idx.ptr
        self.dev = v

    def clear_outputs(self):
        for b in self.ref_out + (self.out or []):
            b.fill(0)

    def result(self):
        """(ref_out, out) as numpy int32 planes (out = ref_out without grain)."""
        def get(b):
            return b.download()[:, :self.cols[id(b)]].astype(np.int32)
        ref = [get(b) for b in self.ref_out]
        out = [get(b) for b in self.out] if self.out is not None else ref
        return ref, out

    def output_pixels(self):
        return int(self.desc.upscaled_width) * int(self.desc.frame_height)

    def hbm_bytes(self):
        """Bytes this frame keeps resident (inputs + outputs, not engine scratch)."""
        bufs = self.src + self.ref_out + (self.out or []) + [self.lf_mi, self.delta_lf, self.cdef_idx]
        bufs += self.tx + [u for u in self.units if u is not None]
        return sum(b.nbytes for b in bufs)

    def free(self):
        bufs = self.src + self.ref_out + (self.out or []) + [self.lf_mi, self.delta_lf, self.cdef_idx]
        bufs += self.tx + [u for u in self.units if u is not None]
        for b in bufs:
            b.free()


class Engine:
    MODE_PLAIN, MODE_BATCH, MODE_GRAPH = 0, 1, 2

    def __init__(self, max_frames=64, tuning="table"):
        """tuning: "table" applies the autotuner's saved table when there is
        one, None keeps the built-in launch shapes, or an abi.Tuning."""
        self.lib = abi.load()
        h = C.c_void_p()
        abi.check(self.lib.oxav1_pf_engine_create(C.byref(h), None, max_frames), "engine_create")
        self.h = h
        self.frames = []
        self.tuning_source = None
        if tuning == "table":
            from . import autotune
            table = autotune.load_table()
            if table is not None:
                self.set_tuning(autotune.to_struct(table))
                self.tuning_source = os.path.basename(autotune.TABLE)
        elif tuning is not None:
            self.set_tuning(tuning)
            self.tuning_source = "given"

    def close(self):
        if self.h:
            self.lib.oxav1_pf_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def prepare(self, frames):
        n = len(frames)
        descs = (abi.FrameDesc * n)(*[f.desc for f in frames])
        devs = (abi.FrameDev * n)(*[f.dev for f in frames])
        abi.check(self.lib.oxav1_pf_engine_prepare(self.h, descs, devs, n), "engine_prepare")
        self.frames = list(frames)

    def run(self, mode):
        abi.check(self.lib.oxav1_pf_engine_run(self.h, mode), "engine_run")

    def sync(self):
        abi.check(self.lib.oxav1_pf_engine_sync(self.h), "engine_sync")

    def set_timing(self, on):
        abi.check(self.lib.oxav1_pf_engine_set_timing(self.h, 1 if on else 0), "set_timing")

    def set_tuning(self, t):
        abi.check(self.lib.oxav1_pf_engine_set_tuning(self.h, C.byref(t)), "set_tuning")

    def tuning(self):
        """(abi.Tuning, CU count, fit) where fit[stage][v] is how many
        workgroups of the stage's register-budget variant v (waves per SIMD
        0 = compiler's choice, 6, 7, 8) fit on one CU."""
        t, cus, fit = abi.Tuning(), C.c_uint32(), ((C.c_uint32 * 4) * 5)()
        abi.check(self.lib.oxav1_pf_engine_get_tuning(self.h, C.byref(t), C.byref(cus), fit), "get_tuning")
        return t, cus.value, [list(r) for r in fit]

    def stage_items(self, stage):
        """Whether a stage has work in the prepared batch."""
        for df in self.frames:
            d = df.desc
            on = {"lf": bool(d.apply_lf), "cdef": bool(d.apply_cdef),
                  "sr": bool(d.use_superres) and d.upscaled_width > d.frame_width,
                  "lr": bool(d.lr.uses_lr), "fg": bool(d.fg.apply_grain)}[stage]
            if on:
                return True
        return False

    def stage_ms(self, stage):
        ms = C.c_double()
        cnt = C.c_uint32()
        abi.check(self.lib.oxav1_pf_engine_stage_ms(self.h, stage.encode(), C.byref(ms), C.byref(cnt)),
                  "stage_ms")
        return ms.value, cnt.value
