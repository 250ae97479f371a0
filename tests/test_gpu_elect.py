"""The encoder's CDEF strength election on the GPU (encoder/cdef_elect.rs):
the per-unit SSD tables of oxav1_pf_cdef_unit_ssd against the oracle's
run_units for every candidate, several frames of different formats in one
launch, u8 and u16 samples; and oxav1_pf_cdef_elect against the oracle's
elect_cdef, field for field, across bit depths, subsamplings, monochrome,
cdef_bits limits and lambdas.  The superres election helpers
(encoder/superres_elect.rs: probe, downscale, upscale of the coded-width
reconstruction) against the oracle's on smooth and noisy planes for every
candidate denominator."""
import ctypes as C

import numpy as np
import pytest

from oxav1 import abi
from oxav1.elect import elect_cdef, synthetic_input

pytestmark = pytest.mark.gpu


def _params(c):
    p = abi.CdefParams()
    p.cdef_damping = c.damping
    if c.plane_set == 0:
        p.y_pri_strength[0], p.y_sec_strength[0] = c.pri, c.sec
    else:
        p.uv_pri_strength[0], p.uv_sec_strength[0] = c.pri, c.sec
    return p


def _ssd_gpu(inps, cands, sample_bytes):
    from oxav1.engine import DevBuffer
    lib = abi.load()
    keep, frames = [], (abi.CdefSsdFrame * len(inps))()
    outs = []
    for f, inp in zip(frames, inps):
        dt = np.uint8 if sample_bytes == 1 else np.uint16
        f.mi_rows, f.mi_cols = inp.mi_rows, inp.mi_cols
        ci = DevBuffer(inp.cdef_idx.shape, np.int8, inp.cdef_idx)
        sk = DevBuffer(inp.skip.shape, np.uint8, inp.skip)
        keep += [ci, sk]
        f.cdef_idx, f.skip = ci.ptr, sk.ptr
        for p in range(inp.num_planes):
            r = DevBuffer(inp.recon[p].shape, dt, inp.recon[p].astype(dt))
            s = DevBuffer(inp.input[p].shape, dt, inp.input[p].astype(dt))
            keep += [r, s]
            f.recon[p], f.input[p] = r.ptr, s.ptr
            f.rows[p], f.cols[p] = inp.recon[p].shape
            f.recon_pitch[p] = f.input_pitch[p] = inp.recon[p].shape[1]
        f.bit_depth = inp.bit_depth
        f.subsampling_x, f.subsampling_y = inp.subsampling_x, inp.subsampling_y
        f.num_planes = inp.num_planes
        f.sample_bytes = sample_bytes
        o = DevBuffer((len(cands), inp.n_units), np.uint64)
        keep.append(o)
        outs.append(o)
        f.ssd = o.ptr
        print(f"ssd frame {inp.recon[0].shape} bd {inp.bit_depth} ss {inp.subsampling_x}{inp.subsampling_y} "
              f"planes {inp.num_planes} sample_bytes {sample_bytes} cands {len(cands)}", flush=True)
    ca = (abi.CdefCand * len(cands))(*cands)
    abi.check(lib.oxav1_pf_cdef_unit_ssd(frames, len(inps), ca, len(cands), None), "cdef_unit_ssd")
    res = [o.download() for o in outs]
    for b in keep:
        b.free()
    return res


FORMATS = [  # (w, h, bit depth, ssx, ssy, mono)
    (136, 72, 8, 1, 1, False),
    (100, 76, 10, 1, 0, False),
    (72, 136, 12, 0, 0, False),
    (96, 64, 8, 1, 1, True),
    (200, 120, 10, 1, 1, False),
]


@pytest.mark.parametrize("sample_bytes", [1, 2])
def test_unit_ssd_matches_oracle(oracle_mod, sample_bytes):
    oo = oracle_mod.load()
    rng = np.random.default_rng(7 + sample_bytes)
    fmts = [f for f in FORMATS if sample_bytes == 2 or f[2] == 8]
    inps = [synthetic_input(40 + i, w, h, bd, sx, sy, mono) for i, (w, h, bd, sx, sy, mono) in enumerate(fmts)]
    for inp in inps:  # an uncoded and an all-skip unit, and a block with idx 1
        inp.cdef_idx[0] = -1
        inp.skip[: 4 * inp.mi_cols] = 1
        if inp.n_units > 2:
            inp.cdef_idx[2] = 1
    cands = [abi.CdefCand(3, 0, 0, 0), abi.CdefCand(3, 1, 0, 0)]
    for _ in range(40):
        cands.append(abi.CdefCand(int(rng.integers(3, 7)), int(rng.integers(0, 2)), int(rng.integers(0, 16)),
                                  int(rng.choice([0, 1, 2, 4]))))
    got = _ssd_gpu(inps, cands, sample_bytes)
    for inp, g in zip(inps, got):
        s = inp.struct()
        for ci, c in enumerate(cands):
            want = (C.c_uint64 * inp.n_units)()
            if c.plane_set == 1 and inp.num_planes == 1:
                continue
            oo.oo_cdef_run_units(C.byref(s), C.byref(_params(c)), c.plane_set, want)
            assert np.array_equal(g[ci], np.array(want, dtype=np.uint64)), (inp.recon[0].shape, ci)


ELECT_CASES = [  # (seed, w, h, bit depth, ssx, ssy, mono, max_bits, lambda)
    (1, 128, 96, 8, 1, 1, False, 0, None),
    (2, 136, 72, 10, 1, 1, False, 3, None),
    (3, 96, 104, 8, 1, 0, False, 2, None),
    (4, 104, 64, 12, 0, 0, False, 1, None),
    (5, 120, 80, 8, 1, 1, True, 3, None),
    (6, 256, 192, 8, 1, 1, False, 3, 0),
    (7, 320, 200, 10, 1, 1, False, 3, 50),
]


def _same(a, b):
    if a is None or b is None:
        return a is None and b is None
    def plan_eq(x, y):
        if x is None or y is None:
            return x is None and y is None
        return x.params == y.params and x.d == y.d and np.array_equal(x.unit_idx, y.unit_idx)
    return plan_eq(a.best, b.best) and plan_eq(a.frame_level, b.frame_level) and a.base_d == b.base_d


@pytest.mark.parametrize("case", ELECT_CASES)
def test_elect_matches_oracle(oracle_mod, case):
    oo = oracle_mod.load()
    seed, w, h, bd, sx, sy, mono, mb, lam = case
    inp = synthetic_input(seed, w, h, bd, sx, sy, mono, lambda_=lam, max_bits=mb)
    if lam == 0:  # exact right half: the per-unit arm wins
        for p in range(inp.num_planes):
            ww = inp.recon[p].shape[1]
            inp.recon[p][:, ww // 2:] = inp.input[p][:, ww // 2:]
    print(f"elect case {case}", flush=True)
    want = elect_cdef(inp, oo, "oo_cdef_elect")
    got = elect_cdef(inp)
    assert _same(got, want), (got, want)


def test_elect_none_cases(oracle_mod):
    inp = synthetic_input(9, 64, 64, 8, 1, 1)
    inp.cdef_idx[:] = -1
    assert elect_cdef(inp) is None
    inp = synthetic_input(9, 64, 64, 8, 1, 1, noise=0)
    for p in range(3):
        inp.recon[p][:] = inp.input[p]  # nothing to gain: the unfiltered frame wins
    assert elect_cdef(inp) is None
    assert elect_cdef(inp, oracle_mod.load(), "oo_cdef_elect") is None


# --- superres election helpers (encoder/superres_elect.rs) ------------------

from oxav1 import elect as E  # noqa: E402

SR_FORMATS = [  # (w, h, bit depth, ssx, ssy, mono)
    (128, 64, 8, 1, 1, False), (200, 72, 10, 1, 0, False), (96, 40, 12, 0, 0, False), (136, 48, 8, 1, 1, True),
    (1920, 1080, 10, 1, 1, False),
]


def _planes(rng, w, h, bd, sx, sy, mono, smooth):
    def one(hh, ww):
        if smooth:
            y, x = np.mgrid[0:hh, 0:ww]
            v = (1 << (bd - 1)) + (1 << (bd - 3)) * np.sin(x / 17.0) * np.cos(y / 11.0)
            return np.clip(v + rng.normal(0, 1, (hh, ww)), 0, (1 << bd) - 1).astype(np.uint16)
        return rng.integers(0, 1 << bd, (hh, ww), dtype=np.uint16)
    pl = [one(h, w)]
    if not mono:
        pl += [one(h >> sy, w >> sx), one(h >> sy, w >> sx)]
    return pl


@pytest.mark.parametrize("fmt", SR_FORMATS)
def test_superres_helpers_match_oracle(oracle_mod, fmt):
    oo = oracle_mod.load()
    w, h, bd, sx, sy, mono = fmt
    rng = np.random.default_rng(w + h)
    print(f"superres helpers {fmt}", flush=True)
    for smooth in (True, False):
        pl = _planes(rng, w, h, bd, sx, sy, mono, smooth)
        assert E.superres_probe(pl[0], bd) == E.superres_probe(pl[0], bd, oo)
        for d in E.candidate_denoms(w) or [16]:
            cw = E.superres_coded_width(w, d)
            if not 0 < cw < w:
                continue
            got = E.downscale_width(pl, sx, sy, cw)
            want = E.downscale_width(pl, sx, sy, cw, oo)
            assert all(np.array_equal(a, b) for a, b in zip(got, want)), d
            # the coded-width reconstruction back to the full width
            rec = [got[0]] + ([np.resize(g, (((h + sy) >> sy), ((cw + sx) >> sx))) for g in got[1:]])
            up = E.upscale_recon(rec, w, (cw + 3) // 4, bd, sx, sy)
            up_o = E.upscale_recon(rec, w, (cw + 3) // 4, bd, sx, sy, oo)
            assert all(np.array_equal(a, b) for a, b in zip(up, up_o)), d


# --- film-grain election helpers (encoder/film_grain_elect.rs) ---------------

FG_FORMATS = [  # (w, h, bit depth, ssx, ssy, mono)
    (150, 98, 8, 1, 1, False), (128, 70, 10, 1, 0, False), (96, 64, 12, 0, 0, False), (120, 66, 8, 1, 1, True),
    (1920, 1080, 10, 1, 1, False),
]


@pytest.mark.parametrize("fmt", FG_FORMATS)
def test_fg_helpers_match_oracle(oracle_mod, fmt):
    from conftest import gen_config
    from oxav1 import Frame
    oo = oracle_mod.load()
    w, h, bd, sx, sy, mono = fmt
    print(f"film-grain helpers {fmt}", flush=True)
    rng = np.random.default_rng(w * h)
    pl = _planes(rng, w, h, bd, sx, sy, mono, True)
    pl = [np.clip(p.astype(np.int64) + rng.normal(0, 3 << (bd - 8), p.shape), 0, (1 << bd) - 1).astype(np.uint16)
          for p in pl]
    den = E.fg_denoise(pl, sx, sy)
    assert all(np.array_equal(a, b) for a, b in zip(den, E.fg_denoise(pl, sx, sy, oo)))
    assert E.fg_noise_estimate(pl, den, sx, sy, bd) == E.fg_noise_estimate(pl, den, sx, sy, bd, oo)
    for p in range(len(pl)):
        assert E.fg_plane_sigma16(pl[p], den[p], bd) == E.fg_plane_sigma16(pl[p], den[p], bd, oo)
    # grain on the reconstruction: planes of ((w + sx) >> sx) x ((h + sy) >> sy)
    f = Frame.generate(gen_config(enable_fg=1, fg_num_y_points=8, fg_ar_lag=int(rng.integers(0, 4)),
                                  fg_chroma_points=0 if mono else 4, fg_overlap=int(rng.integers(0, 2)),
                                  bit_depth=bd, mono=int(mono), subsampling_x=sx, subsampling_y=sy,
                                  noise_amp=8), 0xF6B0 + w)
    rec = [rng.integers(0, 1 << bd, (h, w), dtype=np.uint16)]
    if not mono:
        rec += [rng.integers(0, 1 << bd, ((h + sy) >> sy, (w + sx) >> sx), dtype=np.uint16) for _ in range(2)]
    fg = f.desc.fg
    got = E.fg_apply_grain(fg, rec, bd, sx, sy, f.desc.matrix_coefficients)
    want = E.fg_apply_grain(fg, rec, bd, sx, sy, f.desc.matrix_coefficients, oo)
    assert all(np.array_equal(a, b) for a, b in zip(got, want))
    assert any(not np.array_equal(a, b) for a, b in zip(got, rec))
