"""Host-side frame objects: seeded synthetic frames from the library's
generator as numpy arrays, and the ctypes argument bundles for the five
drop-in entry points (the same structs the CPU oracle takes).
"""
import ctypes as C

import numpy as np

from . import abi


def _arr(ptr, shape, dtype):
    n = int(np.prod(shape))
    if n == 0 or not ptr:
        return np.zeros(shape, dtype=dtype)
    ct = np.ctypeslib.as_ctypes_type(dtype)
    buf = C.cast(ptr, C.POINTER(ct * n)).contents
    return np.frombuffer(buf, dtype=dtype, count=n).reshape(shape).copy()


def planes_struct(arrays):
    """Plane[n] over int32 numpy arrays (which must stay alive)."""
    n = len(arrays)
    ps = (abi.Plane * max(n, 1))()
    for i, a in enumerate(arrays):
        assert a.dtype == np.int32 and a.flags.c_contiguous
        ps[i].samples = a.ctypes.data_as(C.POINTER(C.c_int32))
        ps[i].rows, ps[i].cols = a.shape
    return ps


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct)) if a is not None else C.POINTER(ct)()


def workload_seeds(cfg, n):
    """Seeds of the n synthetic frames of benchmark configuration cfg:
    splitmix64((0x0A1F0000 + cfg) ^ frame_index).  Configuration 3 takes
    its resolution from seed bits 8-9, which are set so the n frames come
    in four equal blocks of 640x360, 1280x720, 1920x1080 and 3840x2160."""
    lib = abi.load()
    base = 0x0A1F0000 + cfg
    seeds = [int(lib.oxav1_gen_splitmix64(base ^ i)) for i in range(n)]
    if cfg == 3:
        block = max(1, (n + 3) // 4)
        seeds = [(s & ~(3 << 8)) | (min(i // block, 3) << 8) for i, s in enumerate(seeds)]
    return seeds


class Frame:
    """One generated frame: padded pre-filter planes + grids + parameters."""

    def __init__(self, gf):
        self.desc = abi.FrameDesc.from_buffer_copy(gf.desc)
        d = self.desc
        np_ = d.num_planes
        self.planes = [_arr(gf.planes[p], (gf.plane_rows[p], gf.plane_cols[p]), np.int32)
                       for p in range(np_)]
        mi = (d.mi_rows, d.mi_cols)
        self.is_intra = _arr(gf.is_intra, mi, np.uint8)
        self.skip = _arr(gf.skip, mi, np.uint8)
        self.ref_frame = _arr(gf.ref_frame, mi, np.uint8)
        self.mode = _arr(gf.mode, mi, np.uint8)
        self.segment_id = _arr(gf.segment_id, mi, np.uint8)
        self.mi_size = _arr(gf.mi_size, mi, np.uint8)
        self.delta_lf = _arr(gf.delta_lf, mi + (4,), np.int8)
        self.lf_mi = _arr(gf.lf_mi, mi, np.uint32)
        self.lf_tx_size = [_arr(gf.lf_tx_size[p], (gf.tx_rows[p], gf.tx_cols[p]), np.uint8)
                           for p in range(3)]
        self.cdef_idx = _arr(gf.cdef_idx, (gf.cdef_rows, gf.cdef_cols), np.int8)
        self.lr_units = []
        for p in range(3):
            n = gf.lr_unit_rows[p] * gf.lr_unit_cols[p]
            if n and gf.lr_units[p]:
                src = C.cast(gf.lr_units[p], C.POINTER(abi.LrUnit * n)).contents
                self.lr_units.append((abi.LrUnit * n).from_buffer_copy(src))
            else:
                self.lr_units.append(None)
        self.lr_unit_shape = [(gf.lr_unit_rows[p], gf.lr_unit_cols[p]) for p in range(3)]

    @classmethod
    def generate(cls, cfg, seed):
        """cfg: a GenConfig or an int benchmark configuration (0..4)."""
        lib = abi.load()
        if isinstance(cfg, int):
            c = abi.GenConfig()
            abi.check(lib.oxav1_gen_default_config(cfg, seed, C.byref(c)), "gen_default_config")
            cfg = c
        out = C.POINTER(abi.GenFrame)()
        abi.check(lib.oxav1_gen_frame_create(C.byref(cfg), seed, C.byref(out)), "gen_frame_create")
        try:
            return cls(out.contents)
        finally:
            lib.oxav1_gen_frame_destroy(out)

    # ---------------- geometry ----------------
    def out_shape(self, p):
        d = self.desc
        sx = d.subsampling_x if p else 0
        sy = d.subsampling_y if p else 0
        return ((d.frame_height + sy) >> sy, (d.upscaled_width + sx) >> sx)

    @property
    def num_planes(self):
        return self.desc.num_planes

    # ---------------- drop-in argument bundles ----------------
    def lf_args(self):
        d = self.desc
        ctx = abi.LfCtx.from_buffer_copy(d.lf)
        g = abi.LfGrids()
        keep = [self.is_intra, self.skip, self.ref_frame, self.mode, self.segment_id,
                self.mi_size, self.delta_lf] + self.lf_tx_size
        g.is_intra = _ptr(self.is_intra, C.c_uint8)
        g.skip = _ptr(self.skip, C.c_uint8)
        g.ref_frame = _ptr(self.ref_frame, C.c_uint8)
        g.mode = _ptr(self.mode, C.c_uint8)
        g.segment_id = _ptr(self.segment_id, C.c_uint8)
        g.delta_lf = _ptr(self.delta_lf, C.c_int8)
        g.mi_size = _ptr(self.mi_size, C.c_uint8)
        for p in range(3):
            g.lf_tx_size[p] = _ptr(self.lf_tx_size[p], C.c_uint8)
        C.memmove(C.addressof(g.seg_feature_active), C.addressof(d.seg_feature_active),
                  C.sizeof(g.seg_feature_active))
        C.memmove(C.addressof(g.seg_feature_data), C.addressof(d.seg_feature_data),
                  C.sizeof(g.seg_feature_data))
        return ctx, g, keep

    def cdef_args(self):
        d = self.desc
        ctx = abi.CdefCtx(mi_rows=d.mi_rows, mi_cols=d.mi_cols, num_planes=d.num_planes,
                          bit_depth=d.bit_depth, subsampling_x=d.subsampling_x,
                          subsampling_y=d.subsampling_y)
        ctx.params = abi.CdefParams.from_buffer_copy(d.cdef)
        g = abi.CdefGrids(cdef_idx=_ptr(self.cdef_idx, C.c_int8), skip=_ptr(self.skip, C.c_uint8))
        return ctx, g, [self.cdef_idx, self.skip]

    def sr_ctx(self):
        d = self.desc
        return abi.SrCtx(use_superres=d.use_superres, num_planes=d.num_planes,
                         bit_depth=d.bit_depth, subsampling_x=d.subsampling_x,
                         subsampling_y=d.subsampling_y, frame_width=d.frame_width,
                         upscaled_width=d.upscaled_width, frame_height=d.frame_height,
                         mi_cols=d.mi_cols)

    def lr_args(self):
        ctx = abi.LrCtx.from_buffer_copy(self.desc.lr)
        g = abi.LrGrids()
        for p in range(3):
            if self.lr_units[p] is not None:
                g.units[p] = C.cast(self.lr_units[p], C.POINTER(abi.LrUnit))
        return ctx, g, list(self.lr_units)
