"""Known-answer tests of the reference's own unit tests, ported to the CPU
oracle (loop_filter.rs:910-1234, cdef.rs:638-1067, superres.rs:524-898,
loop_restoration.rs:998-1704, film_grain.rs:880-1149).  Each test names the
reference test it restates.
"""
import ctypes as C

import numpy as np
import pytest

from oxav1 import abi
from oxav1.frames import planes_struct


@pytest.fixture(scope="module")
def oo(oracle_mod):
    return oracle_mod.load()


# ------------------------------------------------------------------ #
# §7.14 loop_filter.rs                                               #
# ------------------------------------------------------------------ #
def _lf_ctx(levels, mi=4, **kw):
    c = abi.LfCtx()
    for i in range(4):
        c.loop_filter_level[i] = levels[i]
    c.num_planes, c.bit_depth, c.subsampling_x, c.subsampling_y = 3, 8, 1, 1
    c.mi_rows = c.mi_cols = mi
    c.frame_width = c.frame_height = mi * 4
    for k, v in kw.items():
        setattr(c, k, v)
    return c


class _Grids:
    """Constant closures as dense grids (the make_ctx helper, loop_filter.rs:915-956)."""

    def __init__(self, mi, is_intra=1, skip=0, ref=0, mode=0, seg=0, tx=0, misize=0):
        n = mi * mi
        self.a = {k: np.full(n, v, dtype=np.uint8) for k, v in dict(
            is_intra=is_intra, skip=skip, ref_frame=ref, mode=mode, segment_id=seg,
            mi_size=misize).items()}
        self.tx = [np.full(n, tx, dtype=np.uint8) for _ in range(3)]
        self.g = abi.LfGrids()
        for k, v in self.a.items():
            setattr(self.g, k, v.ctypes.data_as(C.POINTER(C.c_uint8)))
        for p in range(3):
            self.g.lf_tx_size[p] = self.tx[p].ctypes.data_as(C.POINTER(C.c_uint8))


def test_strength_returns_zero_when_loop_filter_level_zero(oo):
    c, g = _lf_ctx([0, 0, 0, 0]), _Grids(4)
    assert oo.oo_lf_adaptive_filter_strength(C.byref(c), C.byref(g.g), 0, 0, 0, 0).lvl == 0


def test_strength_basic_table_lvl_32_sharp_0(oo):
    c, g = _lf_ctx([32] * 4), _Grids(4)
    s = oo.oo_lf_adaptive_filter_strength(C.byref(c), C.byref(g.g), 0, 0, 0, 0)
    assert (s.lvl, s.limit, s.blimit, s.thresh) == (32, 32, 100, 2)


def test_strength_intra_ref_delta_lifts_lvl(oo):
    c, g = _lf_ctx([10] * 4, loop_filter_delta_enabled=1), _Grids(4)
    c.loop_filter_ref_deltas[0] = 5
    assert oo.oo_lf_adaptive_filter_strength(C.byref(c), C.byref(g.g), 0, 0, 0, 0).lvl == 15


def test_strength_clipped_at_max_loop_filter(oo):
    c, g = _lf_ctx([60] * 4, loop_filter_delta_enabled=1), _Grids(4)
    c.loop_filter_ref_deltas[0] = 10
    assert oo.oo_lf_adaptive_filter_strength(C.byref(c), C.byref(g.g), 0, 0, 0, 0).lvl == 63


def test_strength_sharpness_and_segment_paths(oo):
    """limit = clamp(lvl >> shift, 1, 9 - sharpness) and the ALT_LF segment term
    (loop_filter.rs:502-523, 543-549), checked against their closed forms."""
    for sharp in range(8):
        for lvl in (1, 7, 20, 33, 63):
            c, g = _lf_ctx([lvl] * 4, loop_filter_sharpness=sharp), _Grids(4)
            s = oo.oo_lf_adaptive_filter_strength(C.byref(c), C.byref(g.g), 0, 0, 0, 0)
            shift = 2 if sharp > 4 else (1 if sharp > 0 else 0)
            lim = min(max(lvl >> shift, 1), 9 - sharp) if sharp else max(lvl >> shift, 1)
            assert (s.limit, s.blimit, s.thresh) == (lim, 2 * (lvl + 2) + lim, lvl >> 4)
    c, g = _lf_ctx([30] * 4), _Grids(4, seg=3)
    g.g.seg_feature_active[3][2] = 1  # SEG_LVL_ALT_LF_Y_H (plane 0, pass 1)
    g.g.seg_feature_data[3][2] = -12
    assert oo.oo_lf_adaptive_filter_strength(C.byref(c), C.byref(g.g), 0, 0, 0, 1).lvl == 18
    assert oo.oo_lf_adaptive_filter_strength(C.byref(c), C.byref(g.g), 0, 0, 0, 0).lvl == 30


def test_filter_size_luma_caps_at_16_chroma_at_8(oo):
    TX_64X64, TX_4X4 = 4, 0
    assert oo.oo_lf_filter_size(TX_64X64, TX_64X64, 0, 0) == 16
    assert oo.oo_lf_filter_size(TX_64X64, TX_64X64, 0, 1) == 8
    assert oo.oo_lf_filter_size(TX_4X4, TX_4X4, 0, 0) == 4


def _plane(arr):
    a = np.ascontiguousarray(arr, dtype=np.int32)
    return a, planes_struct([a])


def test_narrow_filter_idempotent_on_uniform_input(oo):
    a, ps = _plane(np.full((8, 8), 128))
    oo.oo_lf_narrow_filter(ps, 4, 4, 1, 0, 0, 8)
    assert (a == 128).all()


def test_narrow_filter_softens_sharp_step_edge(oo):
    a, ps = _plane(np.where(np.arange(8)[None, :] < 4, 100, 156).repeat(8, 0))
    oo.oo_lf_narrow_filter(ps, 4, 4, 1, 0, 1, 8)
    p0, q0 = a[4, 3], a[4, 4]
    assert p0 > 100 and q0 < 156 and q0 - 128 == 128 - p0


def test_narrow_filter_known_values(oo):
    """Hand-evaluated narrow filter (loop_filter.rs:743-782) on a 100|156 step,
    8-bit (filter4_clamp range [-128, 127]): ps1 = ps0 = -28, qs0 = qs1 = 28.
    hev: filter = clamp(-56) = -56, then clamp(-56 + 3*56) = 112; f1 = 116>>3
    = 14, f2 = 115>>3 = 14 -> p0 = 114, q0 = 142, p1/q1 untouched.
    no hev: filter = clamp(168) = 127; f1 = clamp(131)>>3 = 15, f2 = 15 ->
    p0 = 115, q0 = 141; f = Round2(15, 1) = 8 -> p1 = 108, q1 = 148."""
    a, ps = _plane(np.where(np.arange(8)[None, :] < 4, 100, 156).repeat(8, 0))
    oo.oo_lf_narrow_filter(ps, 4, 4, 1, 0, 1, 8)
    assert list(a[4, 2:6]) == [100, 114, 142, 156]
    b, ps = _plane(np.where(np.arange(8)[None, :] < 4, 100, 156).repeat(8, 0))
    oo.oo_lf_narrow_filter(ps, 4, 4, 1, 0, 0, 8)
    assert list(b[4, 2:6]) == [108, 115, 141, 148]


def test_wide_filter_log2_3_luma_preserves_uniform_input(oo):
    a, ps = _plane(np.full((16, 16), 200))
    oo.oo_lf_wide_filter(ps, 8, 8, 0, 1, 0, 3)
    assert (a == 200).all()


def test_wide_filter_log2_4_luma_preserves_uniform_input(oo):
    a, ps = _plane(np.full((32, 32), 64))
    oo.oo_lf_wide_filter(ps, 16, 16, 0, 1, 0, 4)
    assert (a == 64).all()


def test_wide_filter_matches_spec_weights(oo):
    """Wide filter taps against a direct numpy evaluation of §7.14.6.4 on a ramp."""
    row = np.array([3, 9, 20, 24, 31, 40, 52, 60, 71, 80, 85, 97, 101, 110, 118, 125], dtype=np.int64)
    for log2, plane in ((3, 0), (3, 1), (4, 0)):
        n = 6 if log2 == 4 else (3 if plane == 0 else 2)
        n2 = 0 if (log2 == 3 and plane == 0) else 1
        a, ps = _plane(np.tile(row, (4, 1)))
        x = 8
        oo.oo_lf_wide_filter(ps, x, 1, plane, 1, 0, log2)
        for i in range(-n, n):
            t = sum(row[x + min(max(i + j, -(n + 1)), n)] * (2 if abs(j) <= n2 else 1)
                    for j in range(-n, n + 1))
            assert a[1, x + i] == (t + (1 << (log2 - 1))) >> log2, (log2, plane, i)


def test_filter_mask_flat_region_allows_filter(oo):
    a, ps = _plane(np.full((16, 16), 100))
    m = oo.oo_lf_filter_mask(ps, 8, 8, 0, 16, 32, 8, 1, 0, 8, 8)
    assert m.filter_mask and m.flat_mask and not m.hev_mask


def test_filter_mask_steep_edge_blocks_filter(oo):
    a, ps = _plane(np.where(np.arange(16)[None, :] < 8, 50, 200).repeat(16, 0))
    m = oo.oo_lf_filter_mask(ps, 8, 8, 0, 4, 12, 6, 1, 0, 8, 8)
    assert not m.filter_mask


def test_driver_iterates_correct_edge_count_for_2x2_mi_grid(oo):
    """loop_filter.rs:1183-1233 counts four on-screen edge calls on a 2x2 mi
    grid.  Restated output-exactly: the plane equals the input with the
    pinned narrow filter applied along x = 4 on every row and then along
    y = 4 on every column, and no other sample moves (edges at x = 0 / y = 0
    are off screen)."""
    import lf_edge_cases as lc
    k = lc.case_2x2_grid()
    want = k.expected(oo)
    _, got = k.run(oo.oo_loop_filter_frame)
    assert np.array_equal(got[0], want[0])
    moved = np.argwhere(want[0] != k.planes[0])
    assert len(moved) > 0
    # narrow filter footprint: columns 2..5 (pass 0) or rows 2..5 (pass 1)
    assert all(2 <= x <= 5 or 2 <= y <= 5 for y, x in moved)
    assert len(k.edges) == 4 * 4  # four 4-sample edge segments


def test_get_tx_size_and_residual_size(oo):
    BLOCK_16X16, BLOCK_4X8, BLOCK_128X128 = 6, 1, 15
    assert oo.oo_get_plane_residual_size(BLOCK_16X16, 1, 1, 1) == 3  # BLOCK_8X8
    assert oo.oo_get_plane_residual_size(BLOCK_4X8, 1, 1, 0) == -1  # invalid
    assert oo.oo_get_tx_size(1, 2, BLOCK_128X128, 1, 1) == 3  # 64x64 chroma -> TX_32X32
    assert oo.oo_get_tx_size(0, 7, BLOCK_16X16, 1, 1) == 7


# ------------------------------------------------------------------ #
# §7.15 cdef.rs                                                      #
# ------------------------------------------------------------------ #
def _cdef_ctx(mi=4, np_=1, bd=8, sx=0, sy=0, damping=3):
    c = abi.CdefCtx(mi_rows=mi, mi_cols=mi, num_planes=np_, bit_depth=bd, subsampling_x=sx,
                    subsampling_y=sy)
    c.params.cdef_damping = damping
    return c


def test_constrain(oo):
    assert oo.oo_cdef_constrain(5, 0, 3) == 0 and oo.oo_cdef_constrain(-7, 0, 3) == 0
    assert oo.oo_cdef_constrain(5, 16, 8) == 5 and oo.oo_cdef_constrain(-5, 16, 8) == -5
    assert oo.oo_cdef_constrain(20, 8, 4) == 0


def test_constrain_closed_form(oo):
    for diff in range(-300, 301, 7):
        for thr in (1, 2, 5, 16, 60, 240):
            for damp in (3, 6, 10, 40):
                adj = max(0, damp - (thr.bit_length() - 1))
                ad = abs(diff)
                sh = 0 if adj >= 31 else ad >> adj
                v = min(ad, max(thr - sh, 0))
                assert oo.oo_cdef_constrain(diff, thr, damp) == (v if diff >= 0 else -v)


def _dir(oo, samples, bd=8):
    a = np.ascontiguousarray(samples, dtype=np.int32)
    ps = planes_struct([a])
    c = _cdef_ctx(bd=bd)
    d, v = C.c_int32(), C.c_int32()
    oo.oo_cdef_direction(C.byref(c), ps, 1, 0, 0, C.byref(d), C.byref(v))
    return d.value, v.value


def test_direction_search_uniform_block_returns_zero_var(oo):
    assert _dir(oo, np.full((16, 16), 100)) == (0, 0)


def test_direction_search_stripes(oo):
    y, x = np.mgrid[0:16, 0:16]
    assert _dir(oo, 128 + 10 * y)[0] == 2
    assert _dir(oo, 128 + 10 * x)[0] == 6


def _dir_bruteforce(block, bd):
    """§7.15.2 by brute force: line sums along each direction's projection
    lines, cost = sum over lines of (line sum)^2 * 840 / (line length)."""
    x = (block.astype(np.int64) >> (bd - 8)) - 128
    lines = {
        0: lambda i, j: i + j, 1: lambda i, j: i + j // 2, 2: lambda i, j: i,
        3: lambda i, j: 3 + i - j // 2, 4: lambda i, j: 7 + i - j, 5: lambda i, j: 3 - i // 2 + j,
        6: lambda i, j: j, 7: lambda i, j: i // 2 + j,
    }
    cost = []
    for d in range(8):
        sums, counts = {}, {}
        for i in range(8):
            for j in range(8):
                k = lines[d](i, j)
                sums[k] = sums.get(k, 0) + x[i, j]
                counts[k] = counts.get(k, 0) + 1
        # the spec weights a line of n pixels by 840/n, but caps partial lines of the
        # oblique directions at the full-line weight (n >= 8 counts as 8)
        c = 0
        for k, s in sums.items():
            n = counts[k]
            if d in (1, 3, 5, 7) and 3 <= k <= 7:
                n = 8
            c += s * s * (840 // min(n, 8))
        cost.append(c)
    best = max(range(8), key=lambda d: (cost[d], -d))
    if cost[best] == 0:
        best = 0
    return best, (cost[best] - cost[(best + 4) & 7]) >> 10


def test_direction_search_bruteforce(oo):
    rng = np.random.default_rng(11)
    for trial in range(200):
        bd = [8, 10, 12][trial % 3]
        base = rng.integers(0, 1 << bd, size=(16, 16))
        if trial % 2:
            yy, xx = np.mgrid[0:16, 0:16]
            ang = rng.uniform(0, np.pi)
            base = ((np.cos(ang) * xx + np.sin(ang) * yy) * (4 << (bd - 8)) + base // 4)
            base = np.clip(base, 0, (1 << bd) - 1).astype(np.int64)
        assert _dir(oo, base, bd) == _dir_bruteforce(base[:8, :8], bd), trial


def test_cdef_block_with_idx_neg1_only_copies(oo):
    src = (np.arange(256).reshape(16, 16) % 200).astype(np.int32)
    dst = np.zeros((16, 16), dtype=np.int32)
    c = _cdef_ctx()
    idx = np.full(1, -1, dtype=np.int8)
    skip = np.zeros(16, dtype=np.uint8)
    g = abi.CdefGrids(cdef_idx=idx.ctypes.data_as(C.POINTER(C.c_int8)),
                      skip=skip.ctypes.data_as(C.POINTER(C.c_uint8)))
    oo.oo_cdef_frame(C.byref(c), C.byref(g), planes_struct([src]), 1, planes_struct([dst]), 1)
    assert (dst == src).all()


def test_cdef_block_with_skip_only_copies(oo):
    src = (np.arange(256).reshape(16, 16) * 3 % 200).astype(np.int32)
    dst = np.zeros((16, 16), dtype=np.int32)
    c = _cdef_ctx()
    c.params.y_pri_strength[0], c.params.y_sec_strength[0] = 15, 4
    idx = np.zeros(1, dtype=np.int8)
    skip = np.ones(16, dtype=np.uint8)
    g = abi.CdefGrids(cdef_idx=idx.ctypes.data_as(C.POINTER(C.c_int8)),
                      skip=skip.ctypes.data_as(C.POINTER(C.c_uint8)))
    oo.oo_cdef_frame(C.byref(c), C.byref(g), planes_struct([src]), 1, planes_struct([dst]), 1)
    assert (dst == src).all()


def test_cdef_filter_uniform_block_is_idempotent(oo):
    src = np.full((16, 16), 100, dtype=np.int32)
    dst = src.copy()
    c = _cdef_ctx()
    oo.oo_cdef_filter_block(C.byref(c), planes_struct([src]), 1, planes_struct([dst]), 1, 0, 0, 0,
                            8, 2, 6, 2)
    assert (dst == 100).all()


def test_cdef_filter_block_numpy(oo):
    """§7.15.3 per-sample formula evaluated in numpy on random blocks."""
    rng = np.random.default_rng(5)
    dirs = np.array([[[-1, 1], [-2, 2]], [[0, 1], [-1, 2]], [[0, 1], [0, 2]], [[0, 1], [1, 2]],
                     [[1, 1], [2, 2]], [[1, 0], [2, 1]], [[1, 0], [2, 0]], [[1, 0], [2, -1]]])

    def cons(d, t, damp):
        if t == 0:
            return 0
        adj = max(0, damp - (t.bit_length() - 1))
        v = min(abs(d), max(t - (abs(d) >> adj), 0))
        return v if d >= 0 else -v

    for trial in range(40):
        bd = [8, 10][trial % 2]
        src = rng.integers(0, 1 << bd, size=(16, 16)).astype(np.int32)
        dst = src.copy()
        pri, sec, damp, d = int(rng.integers(0, 16)) << (bd - 8), [0, 1, 2, 4][trial % 4] << (bd - 8), \
            int(rng.integers(3, 7)) + bd - 8, int(rng.integers(0, 8))
        r, cc = 2 * (trial % 2), 2 * ((trial // 2) % 2)
        c = _cdef_ctx(bd=bd)
        oo.oo_cdef_filter_block(C.byref(c), planes_struct([src]), 1, planes_struct([dst]), 1, 0, r,
                                cc, pri, sec, damp, d)
        row = (pri >> (bd - 8)) & 1
        pt = [[4, 2], [3, 3]][row]
        for i in range(8):
            for j in range(8):
                y, x = 4 * r + i, 4 * cc + j
                v0 = int(src[y, x])
                s, mx, mn = 0, v0, v0
                for k in range(2):
                    for sg in (-1, 1):
                        for dd, tap, thr in ((d, pt[k], pri), ((d - 2) & 7, [2, 1][k], sec),
                                             ((d + 2) & 7, [2, 1][k], sec)):
                            yy, xx = y + sg * dirs[dd][k][0], x + sg * dirs[dd][k][1]
                            if 0 <= yy < 16 and 0 <= xx < 16:
                                p = int(src[yy, xx])
                                s += tap * cons(p - v0, thr, damp)
                                mx, mn = max(mx, p), min(mn, p)
                exp = min(max(v0 + ((8 + s - (s < 0)) >> 4), mn), mx)
                assert dst[y, x] == exp, (trial, i, j)


# ------------------------------------------------------------------ #
# §7.16 superres.rs                                                  #
# ------------------------------------------------------------------ #
def _sr(oo, ctx, inputs, shapes):
    o = [np.zeros(s, dtype=np.int32) for s in shapes]
    st = oo.oo_upscale_frame(C.byref(ctx), planes_struct(inputs), len(inputs), planes_struct(o), len(o))
    return st, o


def _srctx(fw, uw, fh, mi_cols, np_=1, sx=0, sy=0, use=1, bd=8):
    return abi.SrCtx(use_superres=use, num_planes=np_, bit_depth=bd, subsampling_x=sx,
                     subsampling_y=sy, frame_width=fw, upscaled_width=uw, frame_height=fh,
                     mi_cols=mi_cols)


def test_use_superres_false_copies(oo):
    st, o = _sr(oo, _srctx(16, 16, 16, 4, use=0), [np.full((16, 16), 137, np.int32)], [(16, 16)])
    assert st == 0 and (o[0] == 137).all()


def test_equal_widths_copies(oo):
    st, o = _sr(oo, _srctx(24, 24, 8, 6), [np.full((8, 24), 99, np.int32)], [(8, 24)])
    assert st == 0 and (o[0] == 99).all()


def test_flat_input_yields_flat_output(oo):
    st, o = _sr(oo, _srctx(8, 16, 4, 4), [np.full((4, 8), 80, np.int32)], [(4, 16)])
    assert st == 0 and (o[0] == 80).all()


def test_upscale_clips_to_bit_depth(oo):
    st, o = _sr(oo, _srctx(8, 24, 1, 6), [np.array([[255, 0] * 4], np.int32)], [(1, 24)])
    assert st == 0 and o[0].min() >= 0 and o[0].max() <= 255


def test_upscaled_smaller_is_rejected(oo):
    st, _ = _sr(oo, _srctx(16, 8, 4, 4), [np.zeros((4, 16), np.int32)], [(4, 8)])
    assert st == abi.ERR_UPSCALED_NOT_LARGER


def test_upscale_plane_geometry_mismatch(oo):
    st, _ = _sr(oo, _srctx(8, 16, 4, 4), [np.zeros((4, 8), np.int32)], [(4, 10)])
    assert st == abi.ERR_PLANE_SHAPE_MISMATCH


def test_chroma_subsampling_halves_widths(oo):
    ins = [np.full((8, 16), 64, np.int32), np.full((4, 8), 128, np.int32), np.full((4, 8), 192, np.int32)]
    st, o = _sr(oo, _srctx(16, 24, 8, 6, np_=3, sx=1, sy=1), ins, [(8, 24), (4, 12), (4, 12)])
    assert st == 0 and (o[0] == 64).all() and (o[1] == 128).all() and (o[2] == 192).all()


def test_upscale_sample_phase_zero_and_edge(oo):
    row = np.arange(16, dtype=np.int32)
    rp = row.ctypes.data_as(C.POINTER(C.c_int32))
    for px in range(3, 13):
        assert oo.oo_upscale_sample(rp, 16, px, 0, 0, 15, 8) == px
    r2 = np.array([10, 0, 0, 0, 0, 0, 0, 0], np.int32)
    assert oo.oo_upscale_sample(r2.ctypes.data_as(C.POINTER(C.c_int32)), 8, 0, 0, 0, 7, 8) == 10


def test_upscale_plane_numpy(oo):
    """upscale_plane against a numpy evaluation of §7.16 (stepX, initialSubpelX,
    8-tap polyphase with the maxX clamp) for every denominator."""
    import json
    import os
    from conftest import ROOT
    filt = np.array(json.load(open(os.path.join(ROOT, "tests", "golden", "spec_tables.json")))
                    ["UPSCALE_FILTER"]).reshape(64, 8)
    rng = np.random.default_rng(3)
    for denom in range(9, 17):
        uw = 96
        fw = (uw * 8 + denom // 2) // denom
        mi_cols = 2 * ((fw + 7) >> 3)
        inp = rng.integers(0, 1024, size=(3, mi_cols * 4)).astype(np.int32)
        st, o = _sr(oo, _srctx(fw, uw, 3, mi_cols, bd=10), [inp], [(3, uw)])
        assert st == 0
        step = ((fw << 14) + uw // 2) // uw
        err = uw * step - (fw << 14)
        num = -((uw - fw) << 13) + uw // 2
        div = int(num / uw)  # truncation toward zero
        init = (div + 128 - int(err / 2)) & 0x3fff
        maxx = mi_cols * 4 - 1
        for x in range(uw):
            sxv = -(1 << 14) + init + x * step
            px, ph = sxv >> 14, (sxv & 0x3fff) >> 8
            acc = sum(int(inp[1, min(max(px + k - 3, 0), maxx)]) * int(filt[ph][k]) for k in range(8))
            assert o[0][1, x] == min(max((acc + 64) >> 7, 0), 1023), (denom, x)


# ------------------------------------------------------------------ #
# §7.17 loop_restoration.rs                                          #
# ------------------------------------------------------------------ #
def test_count_units_in_frame(oo):
    assert [oo.oo_count_units_in_frame(*a) for a in [(64, 256), (64, 100), (64, 0), (256, 240)]] \
        == [4, 2, 1, 1]


def _lrctx(size, uses=1, types=(0, 0, 0), sizes=(256, 256, 256), bd=8, np_=1):
    c = abi.LrCtx(mi_rows=(size + 3) // 4, mi_cols=(size + 3) // 4, num_planes=np_, bit_depth=bd,
                  frame_height=size, upscaled_width=size, uses_lr=uses)
    for p in range(3):
        c.frame_restoration_type[p] = types[p]
        c.loop_restoration_size[p] = sizes[p]
    return c


def test_derive_block_geometry_luma_top_left(oo):
    g = oo.oo_lr_derive_block_geometry(C.byref(_lrctx(64)), 0, 0, 0)
    assert (g.unit_row, g.unit_col, g.x, g.y, g.w, g.h) == (0, 0, 0, 0, 4, 4)
    assert (g.plane_end_x, g.plane_end_y, g.stripe_start_y, g.stripe_end_y) == (63, 63, -8, 55)


def test_get_source_sample_routing(oo):
    def geom(ss, se, pey, y=4):
        return type(oo.oo_lr_derive_block_geometry(C.byref(_lrctx(16)), 0, 0, 0))(
            unit_row=0, unit_col=0, x=0, y=y, w=4, h=4, stripe_start_y=ss, stripe_end_y=se,
            plane_end_x=15, plane_end_y=pey)
    curr, cdef = np.full((16, 16), 11, np.int32), np.full((16, 16), 99, np.int32)
    cp, dp = planes_struct([curr]), planes_struct([cdef])
    g = geom(0, 8, 15)
    assert oo.oo_lr_get_source_sample(cp, 1, dp, 1, 0, 3, 5, C.byref(g)) == 99
    curr[:] = 33
    g = geom(6, 14, 15, 8)
    assert oo.oo_lr_get_source_sample(cp, 1, dp, 1, 0, 3, 2, C.byref(g)) == 33
    curr2, cdef2 = np.full((32, 16), 42, np.int32), np.full((32, 16), 88, np.int32)
    curr2[10] = 7  # y = 12 > stripe_end 8 -> pegged to 10
    g = geom(0, 8, 31, 0)
    assert oo.oo_lr_get_source_sample(planes_struct([curr2]), 1, planes_struct([cdef2]), 1, 0, 3,
                                      12, C.byref(g)) == 7


def test_wiener_coefficients(oo):
    for coeff in ([0, 0, 0], [3, -7, 15], [-5, -23, -17], [10, 8, 46]):
        c = (C.c_int32 * 3)(*coeff)
        f = (C.c_int32 * 7)()
        oo.oo_wiener_coefficients(c, f)
        f = list(f)
        assert sum(f) == 128 and f[:3] == coeff and f[4:] == coeff[::-1]


def _lr_run(oo, curr, cdef, ctx, unit):
    units = (abi.LrUnit * 1)(unit)
    g = abi.LrGrids()
    g.units[0] = C.cast(units, C.POINTER(abi.LrUnit))
    lr = np.zeros_like(cdef)
    oo.oo_loop_restoration_frame(C.byref(ctx), C.byref(g), planes_struct([curr]), 1,
                                 planes_struct([cdef]), 1, planes_struct([lr]), 1)
    return lr


def test_lr_no_op_and_none(oo):
    for uses, t in ((0, 0), (1, 0)):
        lr = _lr_run(oo, np.full((16, 16), 50, np.int32), np.full((16, 16), 200, np.int32),
                     _lrctx(16, uses=uses, types=(t, 0, 0)), abi.LrUnit())
        assert (lr == 200).all()


def test_wiener_identity(oo):
    u = abi.LrUnit(type=1)
    lr = _lr_run(oo, np.full((64, 64), 100, np.int32), np.full((64, 64), 100, np.int32),
                 _lrctx(64, types=(1, 0, 0)), u)
    assert (lr == 100).all()


def test_sgr_uniform_zero_weights(oo):
    u = abi.LrUnit(type=2, sgr_set=0)
    lr = _lr_run(oo, np.full((16, 16), 50, np.int32), np.full((16, 16), 222, np.int32),
                 _lrctx(16, types=(2, 0, 0)), u)
    assert (np.abs(lr[4:12, 4:12] - 222) <= 1).all()


@pytest.mark.parametrize("case", [(0, (0, 128), 137), (10, (128, 0), 200), (14, (0, 128), 88)])
def test_sgr_special_sets(oo, case):
    s, xqd, val = case
    u = abi.LrUnit(type=2, sgr_set=s)
    u.sgr_xqd[0], u.sgr_xqd[1] = xqd
    lr = _lr_run(oo, np.zeros((16, 16), np.int32), np.full((16, 16), val, np.int32),
                 _lrctx(16, types=(2, 0, 0)), u)
    assert (lr == val).all()


def test_sgr_clips_to_range(oo):
    u = abi.LrUnit(type=2, sgr_set=0)
    u.sgr_xqd[0], u.sgr_xqd[1] = -96, 95
    lr = _lr_run(oo, np.zeros((16, 16), np.int32), np.full((16, 16), 128, np.int32),
                 _lrctx(16, types=(2, 0, 0)), u)
    assert lr.min() >= 0 and lr.max() <= 255


def test_box_filter_r0_returns_zeros(oo):
    g0 = oo.oo_lr_derive_block_geometry(C.byref(_lrctx(8)), 0, 0, 0)
    out = (C.c_int32 * 16)(*([7] * 16))
    c = np.full((8, 8), 100, np.int32)
    oo.oo_lr_box_filter(planes_struct([c * 0]), 1, planes_struct([c]), 1, 0, C.byref(g0), 8, 0, 0, 0, out)
    assert list(out) == [0] * 16


# ------------------------------------------------------------------ #
# §7.18.3 film_grain.rs                                              #
# ------------------------------------------------------------------ #
def test_lfsr_draws(oo):
    s = C.c_uint16(1)
    assert oo.oo_fg_random(C.byref(s), 11) == 0x400 and oo.oo_fg_random(C.byref(s), 11) == 0x200
    s = C.c_uint16(0xFF00)
    assert oo.oo_fg_random(C.byref(s), 8) == 0xFF and s.value == 0xFF80


def test_lfsr_bit_model(oo):
    """The LFSR against an independent bit-list model x_{n+16} = x_n ^ x_{n+1} ^ x_{n+3} ^ x_{n+12}."""
    for seed in (1, 0xACE1, 0x1234, 0xFFFF):
        bits = [(seed >> i) & 1 for i in range(16)]
        s = C.c_uint16(seed)
        for _ in range(300):
            bits.append(bits[-16] ^ bits[-15] ^ bits[-13] ^ bits[-4])
            state = sum(b << i for i, b in enumerate(bits[-16:]))
            assert oo.oo_fg_random(C.byref(s), 16) == state


def _fg():
    return abi.FilmGrainParams()


def _lut(oo, fg, np_=3):
    t = (C.c_uint8 * 768)()
    oo.oo_fg_scaling_lookup_init(C.byref(fg), np_, t)
    return np.array(list(t)).reshape(3, 256)


def test_scaling_lookup(oo):
    assert (_lut(oo, _fg()) == 0).all()
    fg = _fg()
    fg.apply_grain, fg.num_y_points = 1, 2
    fg.point_y_value[0], fg.point_y_scaling[0], fg.point_y_value[1], fg.point_y_scaling[1] = 0, 0, 255, 255
    t = _lut(oo, fg, 1)
    assert t[0][0] == 0 and abs(t[0][128] - 128) <= 1 and t[0][255] == 255
    fg.point_y_value[0], fg.point_y_scaling[0], fg.point_y_scaling[1] = 32, 64, 200
    t = _lut(oo, fg, 1)
    assert (t[0][:32] == 64).all() and t[0][255] == 200


def test_scale_lut_interp(oo):
    t = (C.c_uint8 * 768)()
    t[7], t[255] = 42, 99
    assert oo.oo_fg_scale_lut(t, 0, 7, 8) == 42 and oo.oo_fg_scale_lut(t, 0, 255, 8) == 99
    t[2], t[3] = 100, 200
    assert oo.oo_fg_scale_lut(t, 0, 9, 10) == 125


def test_generate_grain_zero_and_first_sample(oo, oracle_mod):
    import json
    import os
    from conftest import ROOT
    gauss = json.load(open(os.path.join(ROOT, "tests", "golden", "spec_tables.json")))["GAUSSIAN_SEQUENCE"]
    L, Cb, Cr = (C.c_int32 * (73 * 82))(), (C.c_int32 * (73 * 82))(), (C.c_int32 * (73 * 82))()
    fg = _fg()
    fg.apply_grain, fg.grain_seed = 1, 12345
    dims = oo.oo_fg_generate_grain(C.byref(fg), 8, 1, 1, L, Cb, Cr)
    assert (dims & 0xFFFF, dims >> 16) == (44, 38)
    assert not any(L) and not any(Cb[:44 * 38]) and not any(Cr[:44 * 38])
    fg.grain_seed, fg.num_y_points, fg.ar_coeff_lag, fg.ar_coeff_shift = 1, 2, 0, 6
    oo.oo_fg_generate_grain(C.byref(fg), 8, 1, 1, L, Cb, Cr)
    s = C.c_uint16(1)
    g0 = gauss[oo.oo_fg_random(C.byref(s), 11)]
    assert L[0] == (g0 + 8) >> 4


def test_generate_grain_ar_numpy(oo):
    """The luma AR recurrence (film_grain.rs:255-289) re-evaluated in numpy."""
    rng = np.random.default_rng(9)
    for lag in range(4):
        fg = _fg()
        fg.apply_grain, fg.grain_seed, fg.num_y_points = 1, int(rng.integers(1, 65535)), 3
        fg.ar_coeff_lag, fg.ar_coeff_shift = lag, int(rng.integers(6, 10))
        for i in range(24):
            fg.ar_coeffs_y_plus_128[i] = int(rng.integers(96, 160))
        L, Cb, Cr = (C.c_int32 * (73 * 82))(), (C.c_int32 * (73 * 82))(), (C.c_int32 * (73 * 82))()
        oo.oo_fg_generate_grain(C.byref(fg), 10, 1, 1, L, Cb, Cr)
        got = np.array(list(L)).reshape(73, 82)
        # white noise from the same generator with lag 0 and AR off
        fg0 = abi.FilmGrainParams.from_buffer_copy(fg)
        fg0.ar_coeff_lag = 0
        oo.oo_fg_generate_grain(C.byref(fg0), 10, 1, 1, L, Cb, Cr)
        w = np.array(list(L)).reshape(73, 82)
        coeffs = [fg.ar_coeffs_y_plus_128[i] - 128 for i in range(24)]
        g = w.copy()
        for y in range(3, 73):
            for x in range(3, 79):
                s, pos = 0, 0
                for dr in range(-lag, 1):
                    for dc in range(-lag, lag + 1):
                        if dr == 0 and dc == 0:
                            break
                        s += g[y + dr, x + dc] * coeffs[pos]
                        pos += 1
                sh = fg.ar_coeff_shift
                g[y, x] = min(max(g[y, x] + ((s + (1 << (sh - 1))) >> sh), -512), 511)
        assert (g == got).all(), lag


def test_film_grain_noops(oo, oracle_mod):
    calls = oracle_mod.calls()
    fg = _fg()
    planes = [np.full((32, 64), 100, np.int32)] * 3
    out = calls.film_grain_synthesis(fg, 8, 3, 0, 0, 0, planes)
    assert all((o == 100).all() for o in out)
    fg.apply_grain, fg.grain_seed, fg.grain_scaling, fg.ar_coeff_shift = 1, 9999, 8, 6
    planes = [np.full((32, 32), 128, np.int32), np.full((32, 32), 64, np.int32), np.full((32, 32), 64, np.int32)]
    out = calls.film_grain_synthesis(fg, 8, 3, 0, 0, 0, planes)
    assert (out[0] == 128).all() and (out[1] == 64).all() and (out[2] == 64).all()


def test_film_grain_luma_ramp_mutates(oo, oracle_mod):
    fg = _fg()
    fg.apply_grain, fg.grain_seed, fg.num_y_points = 1, 7, 2
    fg.point_y_value[0], fg.point_y_scaling[0], fg.point_y_value[1], fg.point_y_scaling[1] = 0, 64, 255, 64
    fg.grain_scaling, fg.ar_coeff_lag, fg.ar_coeff_shift = 8, 0, 6
    out = oracle_mod.calls().film_grain_synthesis(fg, 8, 1, 0, 0, 0, [np.full((64, 64), 128, np.int32)])
    assert (out[0] != 128).any() and out[0].min() >= 0 and out[0].max() <= 255
