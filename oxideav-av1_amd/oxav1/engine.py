"""Device-resident chain engine on torch-allocated HBM buffers.

torch is used only to allocate device memory; every filter runs in the
library's HIP kernels on the engine's own stream.
"""
import ctypes as C

import numpy as np

from . import abi


def _torch():
    import torch
    return torch


class DeviceFrame:
    """One frame's inputs and outputs in HBM, plus its FrameDesc/FrameDev."""

    def __init__(self, frame, device="cuda", fg_out=True):
        torch = _torch()
        self.frame = frame
        d = frame.desc
        self.desc = abi.FrameDesc.from_buffer_copy(d)
        dt = torch.uint8 if d.bit_depth == 8 else torch.int16
        self.sample_dtype = dt
        np_ = d.num_planes

        def to_dev(a, dtype):
            return torch.from_numpy(np.ascontiguousarray(a)).to(device=device, dtype=dtype)

        self.src = [to_dev(frame.planes[p], dt) for p in range(np_)]
        self.ref_out = [torch.zeros(frame.out_shape(p), dtype=dt, device=device) for p in range(np_)]
        self.out = ([torch.zeros(frame.out_shape(p), dtype=dt, device=device) for p in range(np_)]
                    if (fg_out and d.fg.apply_grain) else None)
        self.lf_mi = to_dev(frame.lf_mi.view(np.int32), torch.int32)
        self.delta_lf = to_dev(frame.delta_lf, torch.int8)
        self.tx = [to_dev(frame.lf_tx_size[p], torch.uint8) for p in range(3)]
        self.cdef_idx = to_dev(frame.cdef_idx, torch.int8)
        self.units = []
        for p in range(3):
            u = frame.lr_units[p]
            if u is None:
                self.units.append(None)
            else:
                raw = np.frombuffer(bytes(u), dtype=np.uint8)
                self.units.append(to_dev(raw, torch.uint8))
        v = abi.FrameDev()
        for p in range(np_):
            v.src[p] = self.src[p].data_ptr()
            v.src_pitch[p] = self.src[p].shape[1]
            v.ref_out[p] = self.ref_out[p].data_ptr()
            v.ref_pitch[p] = self.ref_out[p].shape[1]
            if self.out is not None:
                v.out[p] = self.out[p].data_ptr()
                v.out_pitch[p] = self.out[p].shape[1]
        v.lf_mi = self.lf_mi.data_ptr()
        v.delta_lf = self.delta_lf.data_ptr()
        for p in range(3):
            v.lf_tx_size[p] = self.tx[p].data_ptr()
            v.lr_units[p] = self.units[p].data_ptr() if self.units[p] is not None else None
        v.cdef_idx = self.cdef_idx.data_ptr()
        self.dev = v

    def result(self):
        """(ref_out, out) as numpy int32 planes (out = ref_out without grain)."""
        ref = [t.cpu().numpy().astype(np.int32) for t in self.ref_out]
        out = [t.cpu().numpy().astype(np.int32) for t in self.out] if self.out is not None else ref
        return ref, out

    def output_pixels(self):
        return int(self.desc.upscaled_width) * int(self.desc.frame_height)


class Engine:
    MODE_PLAIN, MODE_BATCH, MODE_GRAPH = 0, 1, 2

    def __init__(self, max_frames=64):
        self.lib = abi.load()
        h = C.c_void_p()
        abi.check(self.lib.oxav1_pf_engine_create(C.byref(h), None, max_frames), "engine_create")
        self.h = h
        self.frames = []

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
        _torch().cuda.synchronize()
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

    def stage_ms(self, stage):
        ms = C.c_double()
        cnt = C.c_uint32()
        abi.check(self.lib.oxav1_pf_engine_stage_ms(self.h, stage.encode(), C.byref(ms), C.byref(cnt)),
                  "stage_ms")
        return ms.value, cnt.value

    def stream_handle(self):
        return self.lib.oxav1_pf_engine_stream(self.h)
