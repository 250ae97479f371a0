"""Independent numpy restatements of the AV1 spec arithmetic that the
reference's own tests leave open, checked against the oracle:

  - Wiener filter with non-zero taps (spec 7.17.4; loop_restoration.rs:742-821),
  - self-guided filter and box sums on non-uniform input (spec 7.17.3;
    loop_restoration.rs:485-729),
  - film-grain chroma auto-regression with the luma tap (spec 7.18.3.3;
    film_grain.rs:290-352),
  - film-grain noise stripes, overlap blending and the scaled blend
    (spec 7.18.3.5; film_grain.rs:560-880).

Each model is written from the spec's equations, vectorised where the spec
is a plain loop, and evaluated on inputs where the parts pinned elsewhere
(stripe routing, LFSR, tables) are either identity or taken as given.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

from conftest import ROOT
from oxav1 import abi
from oxav1.frames import planes_struct

TABLES = json.load(open(os.path.join(ROOT, "tests", "golden", "spec_tables.json")))


@pytest.fixture(scope="module")
def oo(oracle_mod):
    return oracle_mod.load()


def r2(x, n):
    """Spec Round2 on int64 arrays (arithmetic shift)."""
    x = np.asarray(x, dtype=np.int64)
    return x if n == 0 else (x + (1 << (n - 1))) >> n


# ------------------------------------------------------------------ #
# Loop restoration                                                   #
# ------------------------------------------------------------------ #
def _lr_oracle(oo, src, bd, unit):
    """Oracle LR on one plane with curr == cdef (so stripe routing reads
    the same samples) and one 64x64 unit covering the plane."""
    h, w = src.shape
    ctx = abi.LrCtx(mi_rows=(h + 3) // 4, mi_cols=(w + 3) // 4, num_planes=1, bit_depth=bd,
                    frame_height=h, upscaled_width=w, uses_lr=1)
    ctx.frame_restoration_type[0] = unit.type
    ctx.loop_restoration_size[0] = 64
    units = (abi.LrUnit * 1)(unit)
    g = abi.LrGrids()
    g.units[0] = C.cast(units, C.POINTER(abi.LrUnit))
    lr = np.zeros_like(src)
    oo.oo_loop_restoration_frame(C.byref(ctx), C.byref(g), planes_struct([src]), 1,
                                 planes_struct([src]), 1, planes_struct([lr]), 1)
    return lr


def wiener_model(src, vco, hco, bd):
    """Spec 7.17.4 on a plane whose rows all lie in one stripe: sources
    clamped to the plane, horizontal pass over rows -3..h+2 rounded by
    InterRound0 and clipped, vertical pass rounded by InterRound1, Clip1."""
    h, w = src.shape
    ir0, ir1 = (5, 9) if bd == 12 else (3, 11)

    def taps(c):
        return [c[0], c[1], c[2], 128 - 2 * (c[0] + c[1] + c[2]), c[2], c[1], c[0]]

    hf, vf = taps(hco), taps(vco)
    p = np.pad(src.astype(np.int64), 3, mode="edge")
    inter = sum(hf[t] * p[:, t:t + w] for t in range(7))
    off = 1 << (bd + 7 - ir0 - 1)
    lim = (1 << (bd + 1 + 7 - ir0)) - 1
    inter = np.clip(r2(inter, ir0), -off, lim - off)
    out = sum(vf[t] * inter[t:t + h, :] for t in range(7))
    return np.clip(r2(out, ir1), 0, (1 << bd) - 1)


def box_model(src, r, eps, bd, pass_):
    """Spec 7.17.3 box filter: A/B over the (2r+1)^2 window (sources clamped
    to the plane) at rows/cols -1..h, then the 3x3 weighted sums (pass 0:
    rows of odd index only, weights 6/5, shift 5 on even rows and 4 on odd
    rows; pass 1: weights 4/3, shift 5)."""
    h, w = src.shape
    n = (2 * r + 1) ** 2
    n2e = n * n * eps
    s = ((1 << 20) + n2e // 2) // n2e
    p = np.pad(src.astype(np.int64), r + 1, mode="edge")
    a = np.zeros((h + 2, w + 2), np.int64)
    b = np.zeros((h + 2, w + 2), np.int64)
    for dy in range(-r, r + 1):
        for dx in range(-r, r + 1):
            win = p[dy + r:dy + r + h + 2, dx + r:dx + r + w + 2]
            a += win * win
            b += win
    a = r2(a, 2 * (bd - 8))
    d = r2(b, bd - 8)
    pp = np.maximum(0, a * n - d * d)
    z = r2(pp * s, 20)
    a2 = np.where(z >= 255, 256, np.where(z == 0, 1, ((z << 8) + z // 2) // (z + 1)))
    one_over_n = ((1 << 12) + n // 2) // n
    bb = r2(((1 << 8) - a2) * b * one_over_n, 12)
    A, B = a2, bb  # index [i + 1, j + 1] for i in -1..h
    f = np.zeros((h, w), np.int64)
    src64 = src.astype(np.int64)
    for i in range(h):
        asum = np.zeros(w, np.int64)
        bsum = np.zeros(w, np.int64)
        shift = 5
        if pass_ == 0 and (i & 1):
            shift = 4
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                if pass_ == 0:
                    wt = (6 if dx == 0 else 5) if ((i + dy) & 1) else 0
                else:
                    wt = 4 if (dx == 0 or dy == 0) else 3
                if wt:
                    asum += wt * A[i + dy + 1, dx + 1:dx + 1 + w]
                    bsum += wt * B[i + dy + 1, dx + 1:dx + 1 + w]
        f[i] = r2(asum * src64[i] + bsum, 8 + shift - 4)
    return f


def sgr_model(src, set_, xqd, bd):
    r0, e0, r1, e1 = np.array(TABLES["SGR_PARAMS"]).reshape(16, 4)[set_].tolist()
    u = src.astype(np.int64) << 4
    w0, w1 = xqd
    w2 = 128 - w0 - w1
    v = w1 * u
    v += w0 * (box_model(src, r0, e0, bd, 0) if r0 else u)
    v += w2 * (box_model(src, r1, e1, bd, 1) if r1 else u)
    return np.clip(r2(v, 11), 0, (1 << bd) - 1)


def _plane(rng, h, w, bd, amp):
    """A smooth ramp plus noise, in range (the spec's frames are post-Clip1)."""
    y, x = np.mgrid[0:h, 0:w]
    base = ((x / w) * 0.6 + (y / h) * 0.4) * ((1 << bd) - 1)
    return np.clip(base + rng.integers(-amp, amp + 1, (h, w)), 0, (1 << bd) - 1).astype(np.int32)


@pytest.mark.parametrize("bd", [8, 10, 12])
def test_wiener_nonzero_taps_numpy(oo, bd):
    rng = np.random.default_rng(100 + bd)
    lo, hi = TABLES["WIENER_TAPS_MIN"], TABLES["WIENER_TAPS_MAX"]
    for h, w in ((40, 52), (56, 88), (21, 37)):
        src = _plane(rng, h, w, bd, 8 << (bd - 8))
        u = abi.LrUnit(type=1)
        vco = [int(rng.integers(lo[k], hi[k] + 1)) for k in range(3)]
        hco = [int(rng.integers(lo[k], hi[k] + 1)) for k in range(3)]
        for k in range(3):
            u.wiener[0][k], u.wiener[1][k] = vco[k], hco[k]
        got = _lr_oracle(oo, src, bd, u)
        want = wiener_model(src, vco, hco, bd)
        assert (got == want).all(), (bd, h, w, vco, hco, np.argwhere(got != want)[:3])


@pytest.mark.parametrize("bd", [8, 10, 12])
def test_sgr_nonuniform_numpy(oo, bd):
    rng = np.random.default_rng(200 + bd)
    lo, hi = TABLES["SGRPROJ_XQD_MIN"], TABLES["SGRPROJ_XQD_MAX"]
    for set_ in range(16):
        h, w = ((36, 44), (56, 64), (23, 30))[set_ % 3]
        src = _plane(rng, h, w, bd, 32 << (bd - 8))
        xqd = [int(rng.integers(lo[k], hi[k] + 1)) for k in range(2)]
        u = abi.LrUnit(type=2, sgr_set=set_)
        u.sgr_xqd[0], u.sgr_xqd[1] = xqd
        got = _lr_oracle(oo, src, bd, u)
        want = sgr_model(src, set_, xqd, bd)
        assert (got == want).all(), (bd, set_, xqd, np.argwhere(got != want)[:3])


# ------------------------------------------------------------------ #
# Film grain                                                         #
# ------------------------------------------------------------------ #
def _grain(oo, fg, bd, sx, sy):
    n = 73 * 82
    L, Cb, Cr = (C.c_int32 * n)(), (C.c_int32 * n)(), (C.c_int32 * n)()
    dims = oo.oo_fg_generate_grain(C.byref(fg), bd, sx, sy, L, Cb, Cr)
    cw, ch = dims & 0xFFFF, dims >> 16
    return (np.array(list(L)).reshape(73, 82), np.array(list(Cb)[:ch * cw]).reshape(ch, cw),
            np.array(list(Cr)[:ch * cw]).reshape(ch, cw))


def _fg_params(rng, bd, lag, sx, sy, *, cfl=0, overlap=1, restricted=0, y_points=6):
    fg = abi.FilmGrainParams()
    fg.apply_grain, fg.grain_seed = 1, int(rng.integers(1, 65536))
    fg.num_y_points = y_points
    ys = np.sort(rng.choice(np.arange(0, 256), size=max(y_points, 1), replace=False))
    for i in range(y_points):
        fg.point_y_value[i], fg.point_y_scaling[i] = int(ys[i]), int(rng.integers(0, 256))
    fg.chroma_scaling_from_luma = cfl
    if not cfl:
        for name in ("cb", "cr"):
            k = int(rng.integers(1, 6))
            vs = np.sort(rng.choice(np.arange(0, 256), size=k, replace=False))
            setattr(fg, f"num_{name}_points", k)
            for i in range(k):
                getattr(fg, f"point_{name}_value")[i] = int(vs[i])
                getattr(fg, f"point_{name}_scaling")[i] = int(rng.integers(0, 256))
    fg.grain_scaling = int(rng.integers(8, 12))
    fg.ar_coeff_lag = lag
    for i in range(24):
        fg.ar_coeffs_y_plus_128[i] = int(rng.integers(0, 256))
    for i in range(25):
        fg.ar_coeffs_cb_plus_128[i] = int(rng.integers(0, 256))
        fg.ar_coeffs_cr_plus_128[i] = int(rng.integers(0, 256))
    fg.ar_coeff_shift = int(rng.integers(6, 10))
    fg.grain_scale_shift = int(rng.integers(0, 4))
    for name in ("cb", "cr"):
        setattr(fg, f"{name}_mult", int(rng.integers(0, 256)))
        setattr(fg, f"{name}_luma_mult", int(rng.integers(0, 256)))
        setattr(fg, f"{name}_offset", int(rng.integers(0, 512)))
    fg.overlap_flag = overlap
    fg.clip_to_restricted_range = restricted
    return fg


@pytest.mark.parametrize("sub", [(1, 1), (1, 0), (0, 0)])
@pytest.mark.parametrize("bd", [8, 10])
def test_chroma_ar_with_luma_tap_numpy(oo, sub, bd):
    """Spec 7.18.3.3 chroma recurrence: the (0,0) tap is the co-located luma
    grain averaged over the subsampling block; white noise from the same
    generator with every chroma AR coefficient zeroed."""
    sx, sy = sub
    rng = np.random.default_rng(300 + 10 * bd + 2 * sx + sy)
    gmin, gmax = -(128 << (bd - 8)), (128 << (bd - 8)) - 1
    for lag in range(4):
        fg = _fg_params(rng, bd, lag, sx, sy)
        L, Cb, Cr = _grain(oo, fg, bd, sx, sy)
        white = abi.FilmGrainParams.from_buffer_copy(fg)
        for i in range(25):
            white.ar_coeffs_cb_plus_128[i] = 128
            white.ar_coeffs_cr_plus_128[i] = 128
        _, Cb0, Cr0 = _grain(oo, white, bd, sx, sy)
        ch, cw = Cb.shape
        sh = fg.ar_coeff_shift
        for got, g0, coef in ((Cb, Cb0, fg.ar_coeffs_cb_plus_128), (Cr, Cr0, fg.ar_coeffs_cr_plus_128)):
            c = [int(coef[i]) - 128 for i in range(25)]
            g = g0.copy()
            for i in range(3, ch):
                for j in range(3, cw - 3):
                    s, pos = 0, 0
                    for dr in range(-lag, 1):
                        for dc in range(-lag, lag + 1):
                            if dr == 0 and dc == 0:
                                ly, lx = ((i - 3) << sy) + 3, ((j - 3) << sx) + 3
                                luma = int(L[ly:ly + sy + 1, lx:lx + sx + 1].sum())
                                s += int(r2(luma, sx + sy)) * c[pos]
                                break
                            s += int(g[i + dr, j + dc]) * c[pos]
                            pos += 1
                        else:
                            continue
                        break
                    g[i, j] = min(max(int(g[i, j]) + int(r2(s, sh)), gmin), gmax)
            assert (g == got).all(), (sub, bd, lag)


def _lfsr(state, bits):
    bit = ((state >> 0) ^ (state >> 1) ^ (state >> 3) ^ (state >> 12)) & 1
    state = (state >> 1) | (bit << 15)
    return (state >> (16 - bits)) & ((1 << bits) - 1), state


def _scale_lut(lut, plane, index, bd):
    if bd == 8:
        return int(lut[plane][index])
    x = index >> (bd - 8)
    rem = index - (x << (bd - 8))
    if x == 255:
        return int(lut[plane][x])
    start, end = int(lut[plane][x]), int(lut[plane][x + 1])
    return start + int(r2((end - start) * rem, bd - 8))


def add_noise_model(fg, bd, planes, sx, sy, mc, tmpl, lut):
    """Spec 7.18.3.5: per 32-row stripe, 34x34 grain blocks at LFSR offsets
    with horizontal overlap (27/17, 17/27 or 23/22 weights), stripes stacked
    with vertical overlap, then the scaled blend (chroma first, from the
    unmodified luma)."""
    H, W = planes[0].shape
    n_pl = len(planes)
    gc = 128 << (bd - 8)
    gmin, gmax = -gc, (256 << (bd - 8)) - 1 - gc
    stripes = []
    for luma_num, _ in enumerate(range(0, (H + 1) // 2, 16)):
        st = fg.grain_seed
        st ^= ((luma_num * 37 + 178) & 255) << 8
        st ^= (luma_num * 173 + 105) & 255
        ns = [np.zeros((34, W + 64), np.int64) for _ in range(n_pl)]
        for x in range(0, (W + 1) // 2, 16):
            rand, st = _lfsr(st, 8)
            ox, oy = rand >> 4, rand & 15
            for p in range(n_pl):
                psx, psy = (sx, sy) if p else (0, 0)
                pox = 6 + ox if psx else 9 + ox * 2
                poy = 6 + oy if psy else 9 + oy * 2
                for i in range(34 >> psy):
                    for j in range(34 >> psx):
                        g = int(tmpl[p][poy + i, pox + j])
                        if psx == 0:
                            if j < 2 and fg.overlap_flag and x > 0:
                                old = int(ns[p][i, x * 2 + j])
                                g = old * 27 + g * 17 if j == 0 else old * 17 + g * 27
                                g = min(max(int(r2(g, 5)), gmin), gmax)
                            ns[p][i, x * 2 + j] = g
                        else:
                            if j == 0 and fg.overlap_flag and x > 0:
                                old = int(ns[p][i, x + j])
                                g = min(max(int(r2(old * 23 + g * 22, 5)), gmin), gmax)
                            ns[p][i, x + j] = g
        stripes.append(ns)
    noise = []
    for p in range(n_pl):
        psx, psy = (sx, sy) if p else (0, 0)
        ph, pw = (H + psy) >> psy, (W + psx) >> psx
        img = np.zeros((ph, pw), np.int64)
        for y in range(ph):
            ln = y >> (5 - psy)
            i = y - (ln << (5 - psy))
            row = stripes[ln][p][i, :pw].copy()
            if fg.overlap_flag and ln > 0:
                old = stripes[ln - 1][p]
                if psy == 0 and i < 2:
                    w0, w1 = (27, 17) if i == 0 else (17, 27)
                    row = np.clip(r2(old[i + 32, :pw] * w0 + row * w1, 5), gmin, gmax)
                elif psy == 1 and i < 1:
                    row = np.clip(r2(old[i + 16, :pw] * 23 + row * 22, 5), gmin, gmax)
            img[y] = row
        noise.append(img)
    if fg.clip_to_restricted_range:
        lo, hi_y = 16 << (bd - 8), 235 << (bd - 8)
        hi_c = hi_y if mc == 0 else 240 << (bd - 8)
    else:
        lo, hi_y, hi_c = 0, (256 << (bd - 8)) - 1, (256 << (bd - 8)) - 1
    shift = fg.grain_scaling
    out = [p.astype(np.int64).copy() for p in planes]
    Y = planes[0].astype(np.int64)
    if n_pl == 3:
        ph, pw = (H + sy) >> sy, (W + sx) >> sx
        for y in range(ph):
            for x in range(pw):
                lx, ly = x << sx, y << sy
                lnx = min(lx + 1, W - 1)
                avg = int(r2(Y[ly, lx] + Y[ly, lnx], 1)) if sx else int(Y[ly, lx])
                for p, npts, mult, lmult, offs in (
                        (1, fg.num_cb_points, fg.cb_mult, fg.cb_luma_mult, fg.cb_offset),
                        (2, fg.num_cr_points, fg.cr_mult, fg.cr_luma_mult, fg.cr_offset)):
                    if not (npts > 0 or fg.chroma_scaling_from_luma):
                        continue
                    orig = int(planes[p][y, x])
                    if fg.chroma_scaling_from_luma:
                        merged = avg
                    else:
                        comb = avg * (lmult - 128) + orig * (mult - 128)
                        merged = min(max((comb >> 6) + ((offs - 256) << (bd - 8)), 0), (1 << bd) - 1)
                    nz = int(r2(_scale_lut(lut, p, merged, bd) * int(noise[p][y, x]), shift))
                    out[p][y, x] = min(max(orig + nz, lo), hi_c)
    if fg.num_y_points > 0:
        for y in range(H):
            for x in range(W):
                orig = int(Y[y, x])
                nz = int(r2(_scale_lut(lut, 0, orig, bd) * int(noise[0][y, x]), shift))
                out[0][y, x] = min(max(orig + nz, lo), hi_y)
    return out


CASES = [
    dict(bd=8, sub=(1, 1), lag=3, overlap=1, size=(70, 100)),
    dict(bd=10, sub=(1, 1), lag=2, overlap=1, size=(99, 67)),
    dict(bd=10, sub=(1, 0), lag=1, overlap=1, size=(64, 90), restricted=1),
    dict(bd=8, sub=(0, 0), lag=0, overlap=1, size=(50, 74), cfl=1),
    dict(bd=12, sub=(1, 1), lag=3, overlap=0, size=(66, 96)),
    dict(bd=10, sub=(0, 0), lag=2, overlap=1, size=(40, 48), restricted=1, mc=0),
    dict(bd=8, sub=(1, 1), lag=1, overlap=1, size=(80, 64), mono=1),
]


@pytest.mark.parametrize("case", CASES, ids=[str(i) for i in range(len(CASES))])
def test_add_noise_synthesis_numpy(oo, oracle_mod, case):
    bd, (sx, sy) = case["bd"], case["sub"]
    rng = np.random.default_rng(400 + len(str(case)))
    fg = _fg_params(rng, bd, case["lag"], sx, sy, cfl=case.get("cfl", 0), overlap=case["overlap"],
                    restricted=case.get("restricted", 0))
    H, W = case["size"]
    n_pl = 1 if case.get("mono") else 3
    planes = [_plane(rng, H, W, bd, 40 << (bd - 8))]
    for _ in range(n_pl - 1):
        planes.append(_plane(rng, (H + sy) >> sy, (W + sx) >> sx, bd, 40 << (bd - 8)))
    mc = case.get("mc", 1)
    got = oracle_mod.calls().film_grain_synthesis(fg, bd, n_pl, sx, sy, mc, [p.copy() for p in planes])
    L, Cb, Cr = _grain(oo, fg, bd, sx, sy)
    t = (C.c_uint8 * 768)()
    oo.oo_fg_scaling_lookup_init(C.byref(fg), n_pl, t)
    lut = np.array(list(t)).reshape(3, 256)
    want = add_noise_model(fg, bd, planes, sx, sy, mc, [L, Cb, Cr], lut)
    for p in range(n_pl):
        d = np.argwhere(got[p] != want[p])
        assert len(d) == 0, (case, p, len(d), d[:3])
    assert any((got[p] != planes[p]).any() for p in range(n_pl))
