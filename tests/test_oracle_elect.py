"""Pins of the oracle's CDEF strength election (encoder/cdef_elect.rs).

The election's arithmetic is the §7.15 filter (pinned elsewhere) plus sums of
squared errors, a fixed candidate search and a rate model.  These tests check
the oracle's election against what fixes it independently of its code:

  - the per-unit SSD tables equal a numpy sum over the oracle's cdef_frame
    output (run_units, cdef_elect.rs:192-240);
  - every returned plan's D equals the SSD of applying that plan through
    cdef_frame with the plan's unit grid (the reference's CdefPlan contract,
    cdef_elect.rs:57-62), and base_d the SSD of the unfiltered frame;
  - the winner's D*256 + lambda*R256 beats the unfiltered frame (which wins
    ties) and is no worse than the frame-level fallback;
  - the frame-level arm (cdef_bits 0) is no worse than a brute-force minimum
    over the coarse candidate grid at dampings 3 and 5, and is elected
    exactly when that minimum beats the unfiltered frame;
  - per-unit plans assign every coded unit the first set of least cost, and
    uncoded units -1;
  - ringing along edges elects the filter and lowers the error (the
    reference's tripwire, tests/cdef_ab.rs:132-140).
"""
import ctypes as C

import numpy as np
import pytest

from oxav1 import abi
from oxav1.elect import elect_cdef, synthetic_input
from oxav1.frames import planes_struct


@pytest.fixture(scope="module")
def oo(oracle_mod):
    return oracle_mod.load()


def unit_ssd(inp, planes, which):
    """Per-64x64-unit SSD of plane set `which` of `planes` against the source."""
    sb_cols = (inp.mi_cols + 15) // 16
    out = np.zeros(inp.n_units, dtype=np.uint64)
    idx = [0] if which == 0 else ([1, 2] if inp.num_planes > 1 else [])
    for p in idx:
        h, w = planes[p].shape
        sy, sx = (inp.subsampling_y, inp.subsampling_x) if p else (0, 0)
        y, x = np.mgrid[0:h, 0:w]
        unit = ((y << sy) // 64) * sb_cols + (x << sx) // 64
        d = planes[p].astype(np.int64) - inp.input[p].astype(np.int64)
        np.add.at(out, unit.ravel(), (d * d).ravel().astype(np.uint64))
    return out


def oracle_cdef(oo, inp, params, unit_idx):
    """cdef_frame over the reconstruction with `params` and a per-unit grid."""
    ctx = abi.CdefCtx()
    ctx.mi_rows, ctx.mi_cols = inp.mi_rows, inp.mi_cols
    ctx.num_planes = inp.num_planes
    ctx.bit_depth = inp.bit_depth
    ctx.subsampling_x, ctx.subsampling_y = inp.subsampling_x, inp.subsampling_y
    ctx.params = params
    ids = np.ascontiguousarray(unit_idx, dtype=np.int8)
    skip = np.ascontiguousarray(inp.skip, dtype=np.uint8)
    g = abi.CdefGrids(ids.ctypes.data_as(C.POINTER(C.c_int8)), skip.ctypes.data_as(C.POINTER(C.c_uint8)))
    src = [p.astype(np.int32) for p in inp.recon]
    dst = [p.copy() for p in src]
    oo.oo_cdef_frame(C.byref(ctx), C.byref(g), planes_struct(src), len(src), planes_struct(dst), len(dst))
    return dst


def params_struct(d):
    p = abi.CdefParams()
    p.cdef_damping, p.cdef_bits = d["cdef_damping"], d["cdef_bits"]
    for i in range(8):
        p.y_pri_strength[i], p.y_sec_strength[i] = d["y_pri"][i], d["y_sec"][i]
        p.uv_pri_strength[i], p.uv_sec_strength[i] = d["uv_pri"][i], d["uv_sec"][i]
    return p


def simple_params(damping, y=(0, 0), uv=(0, 0)):
    p = abi.CdefParams()
    p.cdef_damping = damping
    p.y_pri_strength[0], p.y_sec_strength[0] = y
    p.uv_pri_strength[0], p.uv_sec_strength[0] = uv
    return p


def total_ssd(inp, planes):
    return int(unit_ssd(inp, planes, 0).sum() + unit_ssd(inp, planes, 1).sum())


def rate256(params, inp):
    bits = params["cdef_bits"]
    if bits == 0:
        return 0
    coded = int((np.asarray(inp.cdef_idx) != -1).sum())
    hdr = 12 if inp.num_planes > 1 else 6
    return 256 * (bits * coded + ((1 << bits) - 1) * hdr)


CASES = [  # (seed, w, h, bit depth, ssx, ssy, mono, max_bits)
    (1, 128, 96, 8, 1, 1, False, 0),
    (2, 136, 72, 10, 1, 1, False, 3),
    (3, 96, 104, 8, 1, 0, False, 2),
    (4, 104, 64, 12, 0, 0, False, 1),
    (5, 120, 80, 8, 1, 1, True, 3),
]


@pytest.mark.parametrize("case", CASES)
def test_run_units_tables_match_numpy(oo, case):
    seed, w, h, bd, ssx, ssy, mono, mb = case
    inp = synthetic_input(seed, w, h, bd, ssx, ssy, mono, max_bits=mb)
    rng = np.random.default_rng(seed)
    s = inp.struct()
    for _ in range(3):
        p = simple_params(int(rng.integers(3, 7)), (int(rng.integers(0, 16)), int(rng.choice([0, 1, 2, 4]))),
                          (int(rng.integers(0, 16)), int(rng.choice([0, 1, 2, 4]))))
        ref = oracle_cdef(oo, inp, p, inp.cdef_idx)
        for which in (0, 1):
            got = (C.c_uint64 * inp.n_units)()
            oo.oo_cdef_run_units(C.byref(s), C.byref(p), which, got)
            assert np.array_equal(np.array(got, dtype=np.uint64), unit_ssd(inp, ref, which))


@pytest.mark.parametrize("case", CASES)
def test_plans_are_exact_and_win(oo, case):
    seed, w, h, bd, ssx, ssy, mono, mb = case
    inp = synthetic_input(seed, w, h, bd, ssx, ssy, mono, max_bits=mb)
    r = elect_cdef(inp, oo, "oo_cdef_elect")
    base = total_ssd(inp, [p.astype(np.int32) for p in inp.recon])
    assert r is not None, "ringing content must elect the filter"
    assert r.base_d == base
    for plan in [r.best] + ([r.frame_level] if r.frame_level else []):
        out = oracle_cdef(oo, inp, params_struct(plan.params), plan.unit_idx)
        assert plan.d == total_ssd(inp, out)
        assert np.array_equal(plan.unit_idx == -1, np.asarray(inp.cdef_idx) == -1)
        assert plan.params["cdef_bits"] <= mb
    score = r.best.d * 256 + inp.lambda_ * rate256(r.best.params, inp)
    assert score < base * 256
    if r.frame_level:
        assert r.frame_level.d < base and r.frame_level.params["cdef_bits"] == 0
        assert score <= r.frame_level.d * 256
        assert r.best.params["cdef_bits"] > 0


COARSE = [(0, 0)] + [(p, s) for p in (1, 2, 3, 4, 6, 9, 12, 15) for s in (0, 2)] + [(0, 1), (0, 2), (0, 4)]


@pytest.mark.parametrize("seed,mono,noise", [(11, False, None), (12, True, None), (13, False, 0), (14, False, 40)])
def test_frame_level_vs_brute_force(oo, seed, mono, noise):
    inp = synthetic_input(seed, 72, 64, 8, 1, 1, mono, noise=noise, max_bits=0)
    base = total_ssd(inp, [p.astype(np.int32) for p in inp.recon])
    best = None
    for d in (3, 5):
        ly = min(unit_ssd(inp, oracle_cdef(oo, inp, simple_params(d, y=s), inp.cdef_idx), 0).sum() for s in COARSE)
        luv = 0
        if not mono:
            luv = min(unit_ssd(inp, oracle_cdef(oo, inp, simple_params(d, uv=s), inp.cdef_idx), 1).sum()
                      for s in COARSE)
        best = min(best if best is not None else ly + luv, ly + luv)
    r = elect_cdef(inp, oo, "oo_cdef_elect")
    if best < base:
        assert r is not None and r.best.params["cdef_bits"] == 0 and r.best.d <= best
    else:
        # the coarse grid cannot beat the unfiltered frame; the refinements may
        assert r is None or r.best.d < base


def test_per_unit_choice_is_argmin(oo):
    found = 0
    for seed in (21, 22, 23):
        inp = synthetic_input(seed, 256, 192, 8, 1, 1, lambda_=0, max_bits=3)
        # vary the content per unit: the right half is reconstructed exactly
        # (any filtering there adds error), the left half keeps its noise
        for p in range(3):
            w = inp.recon[p].shape[1]
            inp.recon[p][:, w // 2:] = inp.input[p][:, w // 2:]
        r = elect_cdef(inp, oo, "oo_cdef_elect")
        if r is None or r.best.params["cdef_bits"] == 0:
            continue
        found += 1
        P = r.best.params
        nsets = 1 << P["cdef_bits"]
        s = inp.struct()
        cost = np.zeros((nsets, inp.n_units), dtype=np.uint64)
        for k in range(nsets):
            p = simple_params(P["cdef_damping"], (P["y_pri"][k], P["y_sec"][k]), (P["uv_pri"][k], P["uv_sec"][k]))
            for which in (0, 1):
                t = (C.c_uint64 * inp.n_units)()
                oo.oo_cdef_run_units(C.byref(s), C.byref(p), which, t)
                cost[k] += np.array(t, dtype=np.uint64)
        coded = np.asarray(inp.cdef_idx) != -1
        assert np.array_equal(r.best.unit_idx[coded], np.argmin(cost[:, coded], axis=0))
        assert (r.best.unit_idx[~coded] == -1).all()
        base = unit_ssd(inp, inp.recon, 0) + unit_ssd(inp, inp.recon, 1)
        assert r.best.d == int(cost.min(axis=0)[coded].sum() + base[~coded].sum())
    assert found, "no case elected a per-unit plan"


def test_ringing_tripwire_and_no_coded_units(oo):
    inp = synthetic_input(31, 128, 128, 10, 1, 1, noise=0, max_bits=0)
    r = elect_cdef(inp, oo, "oo_cdef_elect")
    assert r is not None and r.best.d < r.base_d
    assert max(r.best.params["y_pri"][0], r.best.params["y_sec"][0]) > 0
    inp.cdef_idx[:] = -1
    assert elect_cdef(inp, oo, "oo_cdef_elect") is None


# --- superres election helpers (encoder/superres_elect.rs) ------------------

from oxav1 import elect as E  # noqa: E402


@pytest.mark.parametrize("which", ["oracle", "product"])
def test_superres_coded_width_and_candidates(oo, which):
    """The reference's unit tests (superres_elect.rs:240-262); the product's
    two host functions need no device."""
    o = oo if which == "oracle" else None
    assert E.superres_coded_width(128, 16, o) == 64
    assert E.superres_coded_width(96, 12, o) == 64
    assert E.superres_coded_width(96, 16, o) == 48
    for w in range(8, 513, 8):
        for d in E.candidate_denoms(w, o):
            wd = E.superres_coded_width(w, d, o)
            assert wd >= 8 and wd % 8 == 0 and wd < w, (w, d, wd)
    assert E.candidate_denoms(128, o) == [16]
    assert 12 in E.candidate_denoms(96, o)
    # the spec derivation (5.9.8) for every width and denominator
    for w in range(8, 600, 7):
        for d in range(9, 17):
            assert E.superres_coded_width(w, d, o) == (w * 8 + d // 2) // d


def test_downscaler_flat_range_and_ramp(oo):
    # superres_elect.rs:264-276: flat fields stay exact, samples stay in range
    f = [np.full((16, 128), 173, np.uint16)] + [np.full((8, 64), 173, np.uint16)] * 2
    d = E.downscale_width(f, 1, 1, 64, oo)
    assert d[0].shape == (16, 64) and (d[0] == 173).all() and (d[1] == 173).all()
    g = [(np.arange(16 * 128).reshape(16, 128) * 7 % 256).astype(np.uint16)] + [np.zeros((8, 64), np.uint16)] * 2
    assert (E.downscale_width(g, 1, 1, 64, oo)[0] <= 255).all()
    # a linear ramp passes the [1,2,1]/4 low-pass unchanged, so away from the
    # edges each output is the ramp at the centre-aligned position
    # (x + 1/2) * w_src / w_dst - 1/2, taken to 1/64 sample and rounded
    for w_src, w_dst, a in ((128, 96, 3), (100, 72, 5), (64, 40, 1)):
        ramp = (a * np.arange(w_src) + 7).astype(np.uint16)[None, :].repeat(4, 0)
        out = E.downscale_width([ramp], 0, 0, w_dst, oo)[0][0].astype(np.float64)
        x = np.arange(w_dst)
        pos = (x + 0.5) * w_src / w_dst - 0.5
        inner = (pos >= 1) & (pos <= w_src - 2)
        err = np.abs(out[inner] - (a * pos[inner] + 7))
        assert err.max() <= a / 64 + 0.5 + 1e-9


@pytest.mark.parametrize("bd", [8, 10, 12])
def test_superres_probe(oo, bd):
    rng = np.random.default_rng(bd)

    def numpy_probe(y):
        y = y.astype(np.int64)
        s = np.abs(2 * y[:, 1:-1] - y[:, :-2] - y[:, 2:]).sum()
        return (s * 16) // (y.shape[0] * (y.shape[1] - 2)) // (1 << (bd - 8)) < 10

    sc = 1 << (bd - 8)
    assert E.superres_probe(np.full((8, 8), 3 * sc, np.uint16), bd, oo)  # flat: smooth
    cols = np.tile(np.array([0, 60 * sc], np.uint16), (8, 8))  # alternating columns
    assert not E.superres_probe(cols, bd, oo)
    assert not E.superres_probe(np.zeros((8, 2), np.uint16), bd, oo)  # fewer than 3 columns
    for amp in (0, 1, 2, 3, 6, 20):
        y = np.clip(100 * sc + rng.normal(0, amp * sc, (24, 40)), 0, (1 << bd) - 1).astype(np.uint16)
        assert E.superres_probe(y, bd, oo) == numpy_probe(y)


def test_upscale_recon_is_upscale_frame(oo, oracle_mod):
    """upscale_recon is §7.16 on the u16 planes (pinned by the superres
    tests) with the plane shapes of superres_elect.rs:184-195."""
    rng = np.random.default_rng(5)
    for w, h, uw, bd, sx, sy in ((64, 40, 96, 8, 1, 1), (72, 36, 120, 10, 1, 0), (48, 24, 80, 12, 0, 0)):
        pl = [rng.integers(0, 1 << bd, (h, w), dtype=np.uint16)]
        pl += [rng.integers(0, 1 << bd, ((h + sy) >> sy, (w + sx) >> sx), dtype=np.uint16) for _ in range(2)]
        got = E.upscale_recon(pl, uw, (w + 3) // 4, bd, sx, sy, oo)
        ctx = abi.SrCtx(use_superres=1, num_planes=3, bit_depth=bd, subsampling_x=sx, subsampling_y=sy,
                        frame_width=w, upscaled_width=uw, frame_height=h, mi_cols=(w + 3) // 4)
        src = [p.astype(np.int32) for p in pl]
        dst = [np.zeros(g.shape, np.int32) for g in got]
        from oxav1.frames import planes_struct
        assert oo.oo_upscale_frame(C.byref(ctx), planes_struct(src), 3, planes_struct(dst), 3) == 0
        for g, d in zip(got, dst):
            assert np.array_equal(g, d)


# --- film-grain election helpers (encoder/film_grain_elect.rs) ---------------

def _np_denoise(p):
    a = np.pad(p.astype(np.int64), 1, mode="edge")
    k = np.array([[1, 2, 1], [2, 4, 2], [1, 2, 1]])
    s = sum(k[i, j] * a[i:i + p.shape[0], j:j + p.shape[1]] for i in range(3) for j in range(3))
    return ((s + 8) >> 4).astype(np.uint16)


def test_fg_denoise(oo):
    # film_grain_elect.rs:957-964: flat fields stay exact
    f = [np.full((16, 32), 173, np.uint16), np.full((8, 16), 173, np.uint16), np.full((8, 16), 173, np.uint16)]
    d = E.fg_denoise(f, 1, 1, oo)
    assert all(np.array_equal(a, b) for a, b in zip(d, f))
    # a planar ramp is a fixed point of the binomial away from the edges
    y, x = np.mgrid[0:20, 0:30]
    ramp = (3 * x + 5 * y + 11).astype(np.uint16)
    assert np.array_equal(E.fg_denoise([ramp], 0, 0, oo)[0][1:-1, 1:-1], ramp[1:-1, 1:-1])
    # the binomial with edge replication, as a numpy convolution
    rng = np.random.default_rng(3)
    for sx, sy in ((1, 1), (1, 0), (0, 0)):
        pl = [rng.integers(0, 4096, (18, 26), dtype=np.uint16)]
        pl += [rng.integers(0, 4096, (18 >> sy, 26 >> sx), dtype=np.uint16) for _ in range(2)]
        for a, b in zip(E.fg_denoise(pl, sx, sy, oo), pl):
            assert np.array_equal(a, _np_denoise(b))


def _np_sigma16(sq, n):
    return 0 if n == 0 else int(np.round(np.sqrt(sq / n) * 16.0))


@pytest.mark.parametrize("bd", [8, 10])
def test_fg_noise_estimate_and_plane_sigma(oo, bd):
    rng = np.random.default_rng(bd)
    sh = bd - 8
    # a constant offset of k steps: every sigma is 16 |k|
    den = [rng.integers(0, 1 << bd, (40, 64), dtype=np.uint16)] + \
          [rng.integers(0, 1 << bd, (20, 32), dtype=np.uint16) for _ in range(2)]
    k = 3
    inp = [np.clip(d.astype(np.int64) + (k << sh), 0, (1 << bd) - 1).astype(np.uint16) for d in den]
    den = [(i.astype(np.int64) - (k << sh)).astype(np.uint16) for i in inp]  # exact offsets after the clip
    e = E.fg_noise_estimate(inp, den, 1, 1, bd, oo)
    assert e["sigma16"] == 16 * k and e["cb_sigma16"] == 16 * k and e["cr_sigma16"] == 16 * k
    assert all(v == 16 * k for v in e["bin_sigma16"] + e["cb_bin_sigma16"] + e["cr_bin_sigma16"])
    # random noise: the same statistics in numpy (bins of 32 / 64 on the
    # 8-bit scale, sparse bins fall back to the plane's sigma)
    inp = [np.clip(d.astype(np.int64) + rng.normal(0, 6 << sh, d.shape), 0, (1 << bd) - 1).astype(np.uint16)
           for d in den]
    inp[0][:, :8] = 0  # a well-populated darkest bin; several bins stay sparse
    e = E.fg_noise_estimate(inp, den, 1, 1, bd, oo)

    def profile(a, b, nb, bs, floor):
        d = (a.astype(np.int64) - b.astype(np.int64)) >> sh
        sq = d * d
        n = a.size
        g = _np_sigma16(sq.sum(), n)
        bins = np.minimum((a.astype(np.int64) >> sh) >> bs, nb - 1)
        out = []
        for i in range(nb):
            m = bins == i
            out.append(_np_sigma16(sq[m].sum(), m.sum()) if m.sum() >= max(n // 100, floor) else g)
        return g, out
    g, bins = profile(inp[0], den[0], 8, 5, 64)
    assert e["sigma16"] == g and e["bin_sigma16"] == bins
    for p, key in ((1, "cb"), (2, "cr")):
        g, bins = profile(inp[p], den[p], 4, 6, 32)
        assert e[key + "_sigma16"] == g and e[key + "_bin_sigma16"] == bins
    for p in range(3):
        d = (inp[p].astype(np.int64) - den[p].astype(np.int64)) >> sh
        assert E.fg_plane_sigma16(inp[p], den[p], bd, oo) == _np_sigma16((d * d).sum(), d.size)
    assert E.fg_plane_sigma16(inp[0][:0], den[0][:0], bd, oo) == 0


def test_fg_apply_grain_is_synthesis(oo, oracle_mod):
    """apply_grain_to_recon is §7.18.3 (pinned by the film-grain tests) on
    the u16 planes with the shapes of film_grain_elect.rs:847-856."""
    from conftest import gen_config
    from oxav1 import Frame
    f = Frame.generate(gen_config(enable_fg=1, fg_num_y_points=8, fg_ar_lag=2, fg_chroma_points=4, fg_overlap=1,
                                  bit_depth=10, noise_amp=32), 0xF6A1)
    d = f.desc
    planes = [f.planes[p][: f.out_shape(p)[0], : f.out_shape(p)[1]].astype(np.uint16) for p in range(3)]
    got = E.fg_apply_grain(d.fg, planes, 10, d.subsampling_x, d.subsampling_y, d.matrix_coefficients, oo)
    want = oracle_mod.calls().film_grain_synthesis(d.fg, 10, 3, d.subsampling_x, d.subsampling_y,
                                                   d.matrix_coefficients, [p.astype(np.int32) for p in planes])
    assert all(np.array_equal(a, b) for a, b in zip(got, want))
    assert any(not np.array_equal(a, b) for a, b in zip(got, planes))
