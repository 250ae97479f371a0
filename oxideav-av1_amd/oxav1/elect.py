"""Encoder elections on the filter kernels: the CDEF strength election.

`elect_cdef` mirrors the reference's `elect_cdef` (encoder/cdef_elect.rs:136)
over the C ABI entry point `oxav1_pf_cdef_elect`: the committed walker grids
(`cdef_idx` per 64x64 unit, -1 where no idx is coded; `skip` per mi cell),
the source and pre-CDEF reconstruction planes (u16), lambda and the highest
`cdef_bits` the search may propose.  It returns None where the reference
returns None (no coded unit, or the unfiltered frame wins), otherwise a
`CdefElection` with the best plan, the frame-level fallback when present,
and the unfiltered SSD.  `lib` may be any library exporting the same
symbol with the same struct (the tests pass the oracle's).

`synthetic_input` builds an election input of a given shape from a seed:
a source with edges and texture, and a reconstruction with quantisation
noise and ringing along the edges.
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import abi


@dataclass
class CdefPlan:
    params: dict
    unit_idx: np.ndarray
    d: int


@dataclass
class CdefElection:
    best: CdefPlan
    frame_level: CdefPlan = None
    base_d: int = 0


@dataclass
class ElectInput:
    mi_rows: int
    mi_cols: int
    cdef_idx: np.ndarray  # int8, one per 64x64 unit
    skip: np.ndarray      # uint8, one per mi cell
    input: list           # u16 planes
    recon: list           # u16 planes
    bit_depth: int
    subsampling_x: int
    subsampling_y: int
    lambda_: int = 0
    max_bits: int = 0
    keep: list = field(default_factory=list)

    @property
    def num_planes(self):
        return len(self.recon)

    @property
    def n_units(self):
        return ((self.mi_rows + 15) // 16) * ((self.mi_cols + 15) // 16)

    def struct(self):
        s = abi.CdefElectInput()
        s.mi_rows, s.mi_cols = self.mi_rows, self.mi_cols
        arrs = [np.ascontiguousarray(self.cdef_idx, dtype=np.int8), np.ascontiguousarray(self.skip, dtype=np.uint8)]
        s.cdef_idx, s.skip = arrs[0].ctypes.data, arrs[1].ctypes.data
        for p in range(self.num_planes):
            a = np.ascontiguousarray(self.input[p], dtype=np.uint16)
            b = np.ascontiguousarray(self.recon[p], dtype=np.uint16)
            arrs += [a, b]
            s.input[p], s.recon[p] = a.ctypes.data, b.ctypes.data
        h, w = self.recon[0].shape
        s.width, s.height = w, h
        if self.num_planes > 1:
            s.chroma_h, s.chroma_w = self.recon[1].shape
        s.bit_depth = self.bit_depth
        s.subsampling_x, s.subsampling_y = self.subsampling_x, self.subsampling_y
        s.num_planes = self.num_planes
        s.max_bits = self.max_bits
        s.lambda_ = self.lambda_
        self.keep = arrs
        return s


def params_dict(p):
    return {"cdef_damping": p.cdef_damping, "cdef_bits": p.cdef_bits,
            "y_pri": list(p.y_pri_strength), "y_sec": list(p.y_sec_strength),
            "uv_pri": list(p.uv_pri_strength), "uv_sec": list(p.uv_sec_strength)}


def elect_cdef(inp, lib=None, fn="oxav1_pf_cdef_elect"):
    """The election for one frame (see the module docstring)."""
    lib = lib or abi.load()
    s = inp.struct()
    out = abi.CdefElection()
    best_ids = np.zeros(inp.n_units, dtype=np.int8)
    fl_ids = np.zeros(inp.n_units, dtype=np.int8)
    out.best_unit_idx = best_ids.ctypes.data
    out.frame_level_unit_idx = fl_ids.ctypes.data
    st = getattr(lib, fn)(C.byref(s), C.byref(out))
    if st:
        raise abi.Oxav1Error(st, fn)
    if not out.elected:
        return None
    fl = None
    if out.has_frame_level:
        fl = CdefPlan(params_dict(out.frame_level), fl_ids, out.frame_level_d)
    return CdefElection(CdefPlan(params_dict(out.best), best_ids, out.best_d), fl, out.base_d)


def synthetic_input(seed, width, height, bit_depth=8, ssx=1, ssy=1, mono=False, skip_frac=0.25,
                    uncoded_frac=0.1, noise=None, lambda_=None, max_bits=3):
    """An election input with edges, texture, quantisation noise and ringing."""
    rng = np.random.default_rng(seed)
    mx = (1 << bit_depth) - 1
    sc = 1 << (bit_depth - 8)
    noise = 3 * sc if noise is None else noise

    def plane(h, w):
        y, x = np.mgrid[0:h, 0:w]
        base = 96 + 40 * np.sin(x / (7 + rng.uniform(0, 9))) * np.cos(y / (5 + rng.uniform(0, 9)))
        edges = np.zeros((h, w))
        for _ in range(6):
            a, b, c = rng.normal(size=3)
            edges += 50 * ((a * x + b * y + c * 40) > 0)
        src = np.clip((base + edges + rng.normal(0, 2, (h, w))) * sc, 0, mx)
        # ringing: a decaying oscillation after every edge transition
        grad = np.abs(np.diff(edges, axis=1, prepend=edges[:, :1])) > 0
        ring = np.zeros((h, w))
        for k in range(1, 6):
            ring[:, k:] += grad[:, :-k] * (8 * sc) * ((-1) ** k) / k
        rec = np.clip(src + ring + rng.normal(0, noise, (h, w)), 0, mx)
        return src.astype(np.uint16), rec.astype(np.uint16)

    pl = [plane(height, width)]
    if not mono:
        ch, cw = (height + ssy) >> ssy, (width + ssx) >> ssx
        pl += [plane(ch, cw), plane(ch, cw)]
    mi_rows, mi_cols = (height + 3) // 4, (width + 3) // 4
    n_units = ((mi_rows + 15) // 16) * ((mi_cols + 15) // 16)
    cdef_idx = np.where(rng.random(n_units) < uncoded_frac, -1, 0).astype(np.int8)
    skip = (rng.random(mi_rows * mi_cols) < skip_frac).astype(np.uint8)
    if lambda_ is None:
        lambda_ = int(rng.integers(20, 4000)) * sc * sc
    return ElectInput(mi_rows, mi_cols, cdef_idx, skip, [p[0] for p in pl], [p[1] for p in pl], bit_depth,
                      ssx if not mono else 1, ssy if not mono else 1, lambda_, max_bits)


# --- superres election helpers (encoder/superres_elect.rs) -------------------
# Each takes `oracle=None` (the GPU library) or the oracle's ctypes library,
# whose functions of the same names carry an `oo_` prefix.

def _fn(oracle, name):
    return getattr(oracle, "oo_" + name) if oracle is not None else getattr(abi.load(), "oxav1_pf_" + name)


def _ptrs(arrays):
    return (C.c_void_p * 3)(*([a.ctypes.data for a in arrays] + [None] * (3 - len(arrays))))


def superres_coded_width(upscaled_width, denom, oracle=None):
    """superres_coded_width (superres_elect.rs:23-29)."""
    return int(_fn(oracle, "superres_coded_width")(upscaled_width, denom))


def candidate_denoms(width, oracle=None):
    """candidate_denoms (superres_elect.rs:31-45)."""
    out = (C.c_uint32 * 8)()
    n = _fn(oracle, "superres_candidate_denoms")(width, out)
    return list(out[:n])


def superres_probe(y, bit_depth, oracle=None):
    """superres_probe (superres_elect.rs:62-91) on a u16 luma plane."""
    y = np.ascontiguousarray(y, dtype=np.uint16)
    h, w = y.shape
    if oracle is not None:
        return bool(oracle.oo_superres_probe(y.ctypes.data, w, h, bit_depth))
    r = C.c_int32()
    abi.check(abi.load().oxav1_pf_superres_probe(y.ctypes.data, w, h, bit_depth, C.byref(r)), "superres_probe")
    return bool(r.value)


def downscale_width(planes, ssx, ssy, coded_width, oracle=None):
    """downscale_width (superres_elect.rs:93-150): u16 planes -> u16 planes."""
    planes = [np.ascontiguousarray(p, dtype=np.uint16) for p in planes]
    h, w = planes[0].shape
    out = [np.zeros((h, coded_width), np.uint16)]
    if len(planes) > 1:
        out += [np.zeros((h >> ssy, coded_width >> ssx), np.uint16) for _ in range(2)]
    st = _fn(oracle, "superres_downscale")(_ptrs(planes), w, h, ssx, ssy, len(planes), coded_width, _ptrs(out))
    if oracle is None:
        abi.check(st, "superres_downscale")
    return out


def upscale_recon(planes, upscaled_width, mi_cols, bit_depth, ssx, ssy, oracle=None):
    """upscale_recon (superres_elect.rs:152-232): u16 planes at the coded
    width -> u16 planes at upscaled_width."""
    planes = [np.ascontiguousarray(p, dtype=np.uint16) for p in planes]
    h, w = planes[0].shape
    out = [np.zeros((h, upscaled_width), np.uint16)]
    if len(planes) > 1:
        out += [np.zeros(((h + ssy) >> ssy, (upscaled_width + ssx) >> ssx), np.uint16) for _ in range(2)]
    st = _fn(oracle, "superres_upscale_recon")(_ptrs(planes), w, h, upscaled_width, mi_cols, bit_depth, ssx, ssy,
                                               len(planes), _ptrs(out))
    if st:
        raise abi.Oxav1Error(st, "superres_upscale_recon") if oracle is None else RuntimeError(st)
    return out


# --- film-grain election helpers (encoder/film_grain_elect.rs) ---------------

def fg_denoise(planes, ssx, ssy, oracle=None):
    """denoise_frame (film_grain_elect.rs:52-94): u16 planes -> u16 planes."""
    planes = [np.ascontiguousarray(p, dtype=np.uint16) for p in planes]
    h, w = planes[0].shape
    out = [np.zeros_like(p) for p in planes]
    st = _fn(oracle, "fg_denoise")(_ptrs(planes), w, h, ssx, ssy, len(planes), _ptrs(out))
    if oracle is None:
        abi.check(st, "fg_denoise")
    return out


def _estimate_dict(e):
    return {"sigma16": e.sigma16, "bin_sigma16": list(e.bin_sigma16), "cb_sigma16": e.cb_sigma16,
            "cb_bin_sigma16": list(e.cb_bin_sigma16), "cr_sigma16": e.cr_sigma16,
            "cr_bin_sigma16": list(e.cr_bin_sigma16)}


def fg_noise_estimate(planes, denoised, ssx, ssy, bit_depth, oracle=None):
    """noise_estimate (film_grain_elect.rs:97-167) as a dict of its fields."""
    planes = [np.ascontiguousarray(p, dtype=np.uint16) for p in planes]
    denoised = [np.ascontiguousarray(p, dtype=np.uint16) for p in denoised]
    h, w = planes[0].shape
    e = abi.NoiseEstimate()
    st = _fn(oracle, "fg_noise_estimate")(_ptrs(planes), _ptrs(denoised), w, h, ssx, ssy, len(planes), bit_depth,
                                          C.byref(e))
    if oracle is None:
        abi.check(st, "fg_noise_estimate")
    return _estimate_dict(e)


def fg_plane_sigma16(a, b, bit_depth, oracle=None):
    """plane_sigma16 (film_grain_elect.rs:896-910)."""
    a = np.ascontiguousarray(a, dtype=np.uint16).ravel()
    b = np.ascontiguousarray(b, dtype=np.uint16).ravel()
    if oracle is not None:
        return int(oracle.oo_fg_plane_sigma16(a.ctypes.data, b.ctypes.data, a.size, bit_depth))
    r = C.c_uint64()
    abi.check(abi.load().oxav1_pf_fg_plane_sigma16(a.ctypes.data, b.ctypes.data, a.size, bit_depth, C.byref(r)),
              "fg_plane_sigma16")
    return int(r.value)


def fg_apply_grain(fg, planes, bit_depth, ssx, ssy, matrix_coefficients=2, oracle=None):
    """apply_grain_to_recon (film_grain_elect.rs:834-893): the §7.18.3
    synthesis of u16 planes with FilmGrainParams `fg`."""
    planes = [np.ascontiguousarray(p, dtype=np.uint16) for p in planes]
    h, w = planes[0].shape
    out = [np.zeros_like(p) for p in planes]
    st = _fn(oracle, "fg_apply_grain")(C.byref(fg), _ptrs(planes), w, h, bit_depth, ssx, ssy, len(planes),
                                       matrix_coefficients, _ptrs(out))
    if oracle is None:
        abi.check(st, "fg_apply_grain")
    return out


# --- loop-restoration election (encoder/lr_elect.rs) --------------------------

# Default_Restoration_Type_Cdf (cdf.rs:3703)
DEFAULT_RESTORATION_TYPE_CDF = (9413, 22581, 32768, 0)


@dataclass
class LrElectInput:
    input: list           # u16 planes: the source the filters are fitted to
    curr: list            # u16 planes: pre-CDEF reconstruction
    cdef: list            # u16 planes: post-CDEF reconstruction (the LR input)
    bit_depth: int
    subsampling_x: int
    subsampling_y: int
    lambda_: int = 0
    restoration_type_cdf: tuple = DEFAULT_RESTORATION_TYPE_CDF
    mi_rows: int = 0
    mi_cols: int = 0
    keep: list = field(default_factory=list)

    @property
    def num_planes(self):
        return len(self.cdef)

    def struct(self):
        s = abi.LrElectInput()
        arrs = []
        for p in range(self.num_planes):
            for name, src in (("input", self.input), ("curr", self.curr), ("cdef", self.cdef)):
                a = np.ascontiguousarray(src[p], dtype=np.uint16)
                arrs.append(a)
                getattr(s, name)[p] = a.ctypes.data
        h, w = self.cdef[0].shape
        s.width, s.height = w, h
        if self.num_planes > 1:
            s.chroma_h, s.chroma_w = self.cdef[1].shape
        s.bit_depth = self.bit_depth
        s.subsampling_x, s.subsampling_y = self.subsampling_x, self.subsampling_y
        s.num_planes = self.num_planes
        s.mi_rows = self.mi_rows or (h + 3) // 4
        s.mi_cols = self.mi_cols or (w + 3) // 4
        s.lambda_ = self.lambda_
        for i, v in enumerate(self.restoration_type_cdf):
            s.restoration_type_cdf[i] = v
        self.keep = arrs
        return s


def _unit_tuple(u):
    return (u.type, u.sgr_set, tuple(tuple(r) for r in u.wiener), tuple(u.sgr_xqd))


def _cands_dict(c):
    return {"d_none": c.d_none, "d_wiener": c.d_wiener, "d_sgr": list(c.d_sgr),
            "wiener": tuple(tuple(r) for r in c.wiener), "sgr_xqd": [tuple(x) for x in c.sgr_xqd]}


@dataclass
class LrElection:
    elected: bool
    frame_restoration_type: tuple
    uses_lr: int
    uses_chroma_lr: int
    loop_restoration_size: tuple
    unit_rows: tuple
    unit_cols: tuple
    d: int
    d_pre: int
    units: list           # per plane: list of (type, sgr_set, wiener, sgr_xqd) in raster order
    cands: list           # per plane: list of candidate dicts
    raw_units: list = field(default_factory=list)  # the ctypes unit arrays (for apply_lr_plan)

    def struct(self):
        s = abi.LrElection()
        s.elected = int(self.elected)
        for p in range(3):
            s.frame_restoration_type[p] = self.frame_restoration_type[p]
            s.loop_restoration_size[p] = self.loop_restoration_size[p]
            s.unit_rows[p], s.unit_cols[p] = self.unit_rows[p], self.unit_cols[p]
            if p < len(self.raw_units) and self.raw_units[p] is not None:
                s.units[p] = C.addressof(self.raw_units[p])
        s.uses_lr, s.uses_chroma_lr = self.uses_lr, self.uses_chroma_lr
        s.d, s.d_pre = self.d, self.d_pre
        return s


def lr_elect_call(inp, lib=None, fn="oxav1_pf_lr_elect"):
    """The raw call: returns a closure that runs the election into
    preallocated ctypes arrays and returns (out, units, cands)."""
    lib = lib or abi.load()
    s = inp.struct()
    out = abi.LrElection()
    units, cands = [], []
    for p in range(inp.num_planes):
        h, w = inp.cdef[p].shape
        n = max((h + 32) // 64, 1) * max((w + 32) // 64, 1)
        units.append((abi.LrUnit * n)())
        cands.append((abi.LrUnitCands * n)())
        out.units[p] = C.addressof(units[p])
        out.cands[p] = C.addressof(cands[p])
    f = getattr(lib, fn)

    def run():
        st = f(C.byref(s), C.byref(out))
        if st:
            raise abi.Oxav1Error(st, fn)
        return out, units, cands
    return run


def elect_lr(inp, lib=None, fn="oxav1_pf_lr_elect"):
    """The loop-restoration election of one frame (elect_lr,
    encoder/lr_elect.rs:421).  Unlike the reference's Option the result is
    always returned; `elected` is False where the reference returns None.
    `lib` may be the oracle's library (fn="oo_lr_elect")."""
    out, units, cands = lr_elect_call(inp, lib, fn)()
    np_ = inp.num_planes
    return LrElection(
        bool(out.elected), tuple(out.frame_restoration_type), out.uses_lr, out.uses_chroma_lr,
        tuple(out.loop_restoration_size), tuple(out.unit_rows), tuple(out.unit_cols), out.d, out.d_pre,
        [[_unit_tuple(u) for u in units[p][:out.unit_rows[p] * out.unit_cols[p]]] for p in range(np_)],
        [[_cands_dict(c) for c in cands[p][:out.unit_rows[p] * out.unit_cols[p]]] for p in range(np_)],
        units)


def apply_lr_plan(plan, inp, lib=None, fn="oxav1_pf_lr_apply_plan"):
    """apply_lr_plan (lr_elect.rs:829): the plan over copies of inp.cdef;
    returns (restored planes, whole-frame SSD against inp.input)."""
    lib = lib or abi.load()
    s = inp.struct()
    ps = plan.struct()
    recon = [np.ascontiguousarray(c, dtype=np.uint16).copy() for c in inp.cdef]
    ssd = C.c_uint64(0)
    st = getattr(lib, fn)(C.byref(ps), C.byref(s), _ptrs(recon), C.byref(ssd))
    if st:
        raise abi.Oxav1Error(st, fn)
    return recon, ssd.value


def detail_scene(x, y):
    """Wide-band detail with no hard edges (the reference's LR content,
    tests/lr_ab.rs:31-43)."""
    return (120.0 + 34.0 * np.sin(0.13 * x) * np.cos(0.11 * y) + 22.0 * np.sin(0.31 * x + 0.17 * y)
            + 14.0 * np.cos(0.53 * x - 0.29 * y) + 8.0 * np.sin(0.83 * x + 0.71 * y) + 0.15 * x + 0.1 * y)


def synthetic_lr_input(seed, width, height, bit_depth=8, ssx=1, ssy=1, mono=False, lambda_=None, blur=1.0,
                       noise=None, detail=True):
    """An election input: a source of natural-looking detail (or noise
    texture), a pre-CDEF reconstruction that is the source blurred with
    quantisation-like noise, and a post-CDEF plane that smooths it a little
    further."""
    rng = np.random.default_rng(seed)
    mx = (1 << bit_depth) - 1
    sc = 1 << (bit_depth - 8)
    noise = 2.0 * sc if noise is None else noise
    dx, dy = rng.uniform(0, 50, 2)

    def blur3(a, k):
        p = np.pad(a, 1, mode="edge")
        return (k * p[1:-1, 1:-1] + (1 - k) / 4 * (p[:-2, 1:-1] + p[2:, 1:-1] + p[1:-1, :-2] + p[1:-1, 2:]))

    def plane(h, w, step, scale, off):
        y, x = np.mgrid[0:h, 0:w].astype(np.float64)
        if detail:
            base = off + scale * (detail_scene(x * step + dx, y * step + dy) - 120.0)
        else:
            base = off + rng.normal(0, 20, (h, w))
        src = np.clip(np.rint(base * sc), 0, mx)
        rec = blur3(src, 1.0 - 0.6 * blur) + rng.normal(0, noise, (h, w))
        curr = np.clip(np.rint(rec), 0, mx)
        cdef = np.clip(np.rint(blur3(curr, 0.8)), 0, mx)
        return src.astype(np.uint16), curr.astype(np.uint16), cdef.astype(np.uint16)

    pl = [plane(height, width, 1.0, 1.0, 120.0)]
    if not mono:
        ch, cw = (height + ssy) >> ssy, (width + ssx) >> ssx
        pl += [plane(ch, cw, 2.0, 0.4, 128.0), plane(ch, cw, 2.0, -0.35, 128.0)]
    if lambda_ is None:
        lambda_ = int(rng.integers(20, 2000)) * sc * sc
    return LrElectInput([p[0] for p in pl], [p[1] for p in pl], [p[2] for p in pl], bit_depth,
                        ssx if not mono else 1, ssy if not mono else 1, lambda_)
