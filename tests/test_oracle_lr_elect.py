"""Pins of the oracle's loop-restoration election (encoder/lr_elect.rs) and
of the rate model it prices units with (encoder/symbol_writer.rs counting
mode, encoder/loop_restoration_write.rs:95-203).

The election's arithmetic is the §7.17 filter (pinned elsewhere), sums of
squared errors, two least-squares fits and a rate model.  These tests check
the oracle against what fixes each part independently of its code:

  - every unit's rects partition the plane, and each sample lies in the
    unit §7.17.1 assigns it (unitRow from (lumaY + 8) >> subY, clamped);
  - the per-unit candidate SSDs equal numpy sums over the oracle's
    loop_restoration_frame output with every unit set to that candidate;
  - the self-guided weights equal a numpy least-squares solve of the
    (128, 0) / (0, 128) probe deltas, rounded half away from zero and
    constrained as §5.11.58 constrains them;
  - every step of the Wiener fit equals a numpy lstsq over the explicitly
    built basis (m_j from the fixed 7-tap filter, the symmetric basis, the
    target src - m_3), clamped to the tap ranges, and the quantised taps
    round the last steps;
  - log2_frac256 meets the reference's anchor test and lies within one
    1/256 of 256*log2(x/2^15); symbol costs match -256*log2(p); the
    reference's sub-bit cost test; every subexp code decodes back through
    the §5.9.28 decoder and costs one bit per bool;
  - every elected unit is the first argmin of D*256 + lambda*R over
    {none, Wiener, the first self-guided set of least SSD}, priced from the
    running references; lambda = 0 elects the least SSD, a huge lambda
    elects nothing; d and d_pre are the sums of the elected and unfiltered
    SSDs, and apply_lr_plan reproduces d (the reference's debug_assert);
  - detail content under a coarse reconstruction elects restoration (the
    reference's tripwire, tests/lr_ab.rs:229-254).
"""
import ctypes as C

import numpy as np
import pytest

from oxav1 import abi
from oxav1.elect import DEFAULT_RESTORATION_TYPE_CDF, apply_lr_plan, elect_lr, synthetic_lr_input
from oxav1.frames import planes_struct

WIENER_MIN = [-5, -23, -17]
WIENER_MAX = [10, 8, 46]
WIENER_MID = [3, -7, 15]
XQD_MIN = [-96, -32]
XQD_MAX = [31, 95]
SGR_R = [(2, 1)] * 10 + [(0, 1)] * 4 + [(2, 0)] * 2   # (r0, r1) of the 16 Sgr_Params sets (§7.17.3)

CASES = [  # seed, width, height, bit depth, ssx, ssy, mono, lambda
    (1, 100, 76, 8, 1, 1, False, 40),
    (2, 130, 98, 10, 1, 0, False, 300),
    (3, 96, 100, 12, 0, 0, True, 800),
]


@pytest.fixture(scope="module")
def oo(oracle_mod):
    return oracle_mod.load()


@pytest.fixture(scope="module", params=CASES, ids=lambda c: f"{c[1]}x{c[2]}-{c[3]}bit-{c[4]}{c[5]}{'-mono' if c[6] else ''}")
def case(request, oo):
    seed, w, h, bd, sx, sy, mono, lam = request.param
    inp = synthetic_lr_input(seed, w, h, bd, sx, sy, mono, lambda_=lam)
    return inp, elect_lr(inp, oo, "oo_lr_elect")


def _planes(inp):
    return [c.shape for c in inp.cdef]


def unit_map(inp, p, urows, ucols):
    """§7.17.1: the unit of every sample of plane p."""
    h, w = inp.cdef[p].shape
    sy, sx = (inp.subsampling_y, inp.subsampling_x) if p else (0, 0)
    y, x = np.mgrid[0:h, 0:w]
    # luma row of the sample's 4x4 block: (y << sy) rounded down to the block
    luma_y = ((y << sy) // 4) * 4
    ur = np.minimum(urows - 1, ((luma_y + 8) >> sy) // 64)
    luma_x = ((x << sx) // 4) * 4
    uc = np.minimum(ucols - 1, (luma_x >> sx) // 64)
    return ur * ucols + uc


def lr_frame(oo, inp, grids, frt=None):
    """oo_loop_restoration_frame with 64x64 units, grids[p] a LrUnit array."""
    h, w = inp.cdef[0].shape
    np_ = inp.num_planes
    ctx = abi.LrCtx(mi_rows=(h + 3) // 4, mi_cols=(w + 3) // 4, num_planes=np_, bit_depth=inp.bit_depth,
                    subsampling_x=inp.subsampling_x, subsampling_y=inp.subsampling_y, frame_height=h,
                    upscaled_width=w, uses_lr=1)
    g = abi.LrGrids()
    for p in range(np_):
        ctx.frame_restoration_type[p] = abi.RESTORE_SWITCHABLE if frt is None else frt[p]
        ctx.loop_restoration_size[p] = 64
        g.units[p] = C.cast(grids[p], C.POINTER(abi.LrUnit))
    curr = [np.ascontiguousarray(c, dtype=np.int32) for c in inp.curr]
    cdef = [np.ascontiguousarray(c, dtype=np.int32) for c in inp.cdef]
    lr = [np.zeros_like(c) for c in cdef]
    oo.oo_loop_restoration_frame(C.byref(ctx), C.byref(g), planes_struct(curr), np_, planes_struct(cdef), np_,
                                 planes_struct(lr), np_)
    return lr


def grid_of(res, p, make):
    n = res.unit_rows[p] * res.unit_cols[p]
    arr = (abi.LrUnit * n)()
    for k in range(n):
        arr[k] = make(k)
    return arr


def unit_sums(values, umap, n):
    out = np.zeros(n, dtype=np.int64)
    np.add.at(out, umap.ravel(), values.ravel().astype(np.int64))
    return out


def test_unit_rects_partition_the_plane(case, oo):
    inp, res = case
    s = inp.struct()
    for p in range(inp.num_planes):
        h, w = inp.cdef[p].shape
        umap = unit_map(inp, p, res.unit_rows[p], res.unit_cols[p])
        seen = np.full((h, w), -1)
        for k in range(res.unit_rows[p] * res.unit_cols[p]):
            ur, uc = divmod(k, res.unit_cols[p])
            rects = (C.c_uint32 * (4 * 4096))()
            n = oo.oo_lr_unit_rects(C.byref(s), p, ur, uc, rects, 4096)
            last = (-1, -1)
            for i in range(n):
                x, y, rw, rh = rects[4 * i:4 * i + 4]
                assert (y, x) > last, "blocks are pushed in raster order"
                last = (y, x)
                assert (seen[y:y + rh, x:x + rw] == -1).all()
                seen[y:y + rh, x:x + rw] = k
        assert (seen == umap).all()


def test_candidate_ssd_tables(case, oo):
    inp, res = case
    for p in range(inp.num_planes):
        nu = res.unit_rows[p] * res.unit_cols[p]
        umap = unit_map(inp, p, res.unit_rows[p], res.unit_cols[p])
        src = inp.input[p].astype(np.int64)
        d0 = unit_sums((inp.cdef[p].astype(np.int64) - src) ** 2, umap, nu)
        assert [c["d_none"] for c in res.cands[p]] == list(d0)

        def run(make):
            grids = [grid_of(res, q, make if q == p else (lambda k: abi.LrUnit())) for q in range(inp.num_planes)]
            out = lr_frame(oo, inp, grids)[p]
            return list(unit_sums((out.astype(np.int64) - src) ** 2, umap, nu))

        def wiener(k):
            u = abi.LrUnit(type=abi.RESTORE_WIENER)
            for a in range(2):
                for b in range(3):
                    u.wiener[a][b] = res.cands[p][k]["wiener"][a][b]
            return u

        assert run(wiener) == [c["d_wiener"] for c in res.cands[p]]
        for s in range(16):
            def sgr(k, s=s):
                u = abi.LrUnit(type=abi.RESTORE_SGRPROJ, sgr_set=s)
                u.sgr_xqd[0], u.sgr_xqd[1] = res.cands[p][k]["sgr_xqd"][s]
                return u
            assert run(sgr) == [c["d_sgr"][s] for c in res.cands[p]], f"set {s}"


def round_half_away(x):
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


def test_sgr_weights_are_probe_least_squares(case, oo):
    inp, res = case
    ties = 0
    for p in range(inp.num_planes):
        nu = res.unit_rows[p] * res.unit_cols[p]
        umap = unit_map(inp, p, res.unit_rows[p], res.unit_cols[p])
        cdef = inp.cdef[p].astype(np.int64)
        t = 128 * (inp.input[p].astype(np.int64) - cdef)
        for s in range(16):
            r0, r1 = SGR_R[s]

            def probe(xqd):
                def mk(k):
                    u = abi.LrUnit(type=abi.RESTORE_SGRPROJ, sgr_set=s)
                    u.sgr_xqd[0], u.sgr_xqd[1] = xqd
                    return u
                grids = [grid_of(res, q, mk if q == p else (lambda k: abi.LrUnit())) for q in range(inp.num_planes)]
                return lr_frame(oo, inp, grids)[p].astype(np.int64) - cdef

            a0 = probe((128, 0)) if r0 else np.zeros_like(cdef)
            a1 = probe((0, 128)) if r1 else np.zeros_like(cdef)
            s00, s01, s11 = (unit_sums(a0 * a0, umap, nu), unit_sums(a0 * a1, umap, nu), unit_sums(a1 * a1, umap, nu))
            b0, b1 = unit_sums(a0 * t, umap, nu), unit_sums(a1 * t, umap, nu)
            for k in range(nu):
                xq = [0.0, 0.0]
                if r0 and r1:
                    m = np.array([[s00[k], s01[k]], [s01[k], s11[k]]], dtype=np.float64)
                    if abs(np.linalg.det(m)) > 1e-6:
                        xq = list(np.linalg.solve(m, np.array([b0[k], b1[k]], dtype=np.float64)))
                elif r0 and s00[k] > 0:
                    xq[0] = b0[k] / s00[k]
                elif r1 and s11[k] > 0:
                    xq[1] = b1[k] / s11[k]
                if any(abs(abs(v - np.floor(v)) - 0.5) < 1e-7 for v in xq):
                    ties += 1
                    continue
                q0 = int(np.clip(round_half_away(xq[0]), XQD_MIN[0], XQD_MAX[0]))
                q1 = int(np.clip(round_half_away(xq[1]), XQD_MIN[1], XQD_MAX[1]))
                if not r0:
                    q0 = 0
                if not r1:
                    q1 = int(np.clip(128 - q0, XQD_MIN[1], XQD_MAX[1]))
                assert tuple(res.cands[p][k]["sgr_xqd"][s]) == (q0, q1), (p, k, s, xq)
    assert ties <= 2


def taps7(t):
    c = 128.0 - 2.0 * (t[0] + t[1] + t[2])
    return [t[0], t[1], t[2], c, t[2], t[1], t[0]]


def fit_step(cdef, src, pts, fixed, direction, fc):
    """One alternating least-squares step: the basis from the fixed taps'
    filter, solved by numpy lstsq and clamped to the tap ranges."""
    h, w = cdef.shape
    pad = np.pad(cdef.astype(np.float64), 6, mode="edge")
    xs, ys = pts[:, 0], pts[:, 1]
    m = np.zeros((len(pts), 7))
    for j in range(7):
        acc = np.zeros(len(pts))
        for k in range(7):
            if direction == 0:
                acc += fixed[k] * pad[ys + k - 3 + 6, xs + j - 3 + 6]
            else:
                acc += fixed[k] * pad[ys + j - 3 + 6, xs + k - 3 + 6]
        m[:, j] = acc / 128.0
    target = src[ys, xs].astype(np.float64) - m[:, 3]
    basis = np.stack([(m[:, i] + m[:, 6 - i] - 2.0 * m[:, 3]) / 128.0 for i in range(3)], axis=1)[:, fc:]
    sol, _, rank, _ = np.linalg.lstsq(basis, target, rcond=None)
    if rank < 3 - fc:
        return None
    return [float(np.clip(sol[i - fc], WIENER_MIN[i], WIENER_MAX[i])) for i in range(fc, 3)]


def test_wiener_fit_steps_are_least_squares(case, oo):
    inp, res = case
    s = inp.struct()
    checked = 0
    for p in range(inp.num_planes):
        fc = 1 if p else 0
        h, w = inp.cdef[p].shape
        cdef = np.ascontiguousarray(inp.cdef[p], dtype=np.int32)
        src = np.ascontiguousarray(inp.input[p], dtype=np.uint16)
        for k in range(res.unit_rows[p] * res.unit_cols[p]):
            ur, uc = divmod(k, res.unit_cols[p])
            rects = (C.c_uint32 * (4 * 4096))()
            n = oo.oo_lr_unit_rects(C.byref(s), p, ur, uc, rects, 4096)
            taps = (C.c_int32 * 6)()
            trace = (C.c_double * 12)()
            oo.oo_lr_fit_wiener(cdef.ctypes.data_as(C.POINTER(C.c_int32)), src.ctypes.data_as(C.POINTER(C.c_uint16)),
                                w, h, rects, n, fc, taps, trace)
            assert [list(taps[0:3]), list(taps[3:6])] == [list(r) for r in res.cands[p][k]["wiener"]]
            pts = []
            for i in range(n):
                x, y, rw, rh = rects[4 * i:4 * i + 4]
                pts += [(xx, yy) for yy in range(y, y + rh) for xx in range(x, x + rw)]
            pts = np.array(pts)
            vt = [0.0 if (fc and i == 0) else float(WIENER_MID[i]) for i in range(3)]
            ht = list(vt)
            steps = [list(trace[3 * i:3 * i + 3]) for i in range(4)]
            for step in range(4):
                direction = step & 1
                fixed = taps7(vt if direction == 0 else ht)
                sol = fit_step(inp.cdef[p], inp.input[p], pts, fixed, direction, fc)
                if sol is not None:
                    np.testing.assert_allclose(steps[step][fc:], sol, rtol=1e-6, atol=1e-6)
                    checked += 1
                if direction == 0:
                    ht = steps[step]
                else:
                    vt = steps[step]
            q = [[int(np.clip(round_half_away(v), WIENER_MIN[i], WIENER_MAX[i])) for i, v in enumerate(t)]
                 for t in (steps[3], steps[2])]
            if fc:
                q[0][0] = q[1][0] = 0
            assert q == [list(taps[0:3]), list(taps[3:6])]
    assert checked > 0


# ---- rate model ---------------------------------------------------------

def test_log2_frac256(oo):
    f = oo.oo_log2_frac256
    assert f(1 << 15) == 0 and f((1 << 16) - 1) == 255 and f(46341) == 128
    xs = np.arange(1 << 15, 1 << 16)
    vals = np.array([f(int(x)) for x in xs])
    assert (np.diff(vals) >= 0).all()
    exact = 256 * np.log2(xs / 32768.0)
    assert (vals <= exact + 1e-9).all() and (exact - vals < 1).all()


def test_symbol_costs_follow_probabilities(oo):
    cdf = (C.c_uint16 * 4)(*DEFAULT_RESTORATION_TYPE_CDF)
    cum = [0] + list(DEFAULT_RESTORATION_TYPE_CDF[:3])
    for s in range(3):
        p = (cum[s + 1] - cum[s]) / 32768.0
        cost = oo.oo_cw_cost(0x8000, 0, s, 3, 0, 0, 0, cdf)
        assert abs(cost - (-256 * np.log2(p))) <= 2, (s, cost)
    # the reference's sub-bit test (symbol_writer.rs:1234-1258)
    cdf2 = (C.c_uint16 * 3)(4096, 32768, 0)
    likely = oo.oo_cw_symbol_run(1 << 15, 1, cdf2, 2, 64) / 64 / 256
    rare = oo.oo_cw_symbol_run(1 << 15, 0, cdf2, 2, 64) / 64 / 256
    assert 0.05 < likely < 0.5 and 2.0 < rare < 4.0


class BitReader:
    def __init__(self, bits):
        self.bits, self.i = list(bits), 0

    def L(self, n):  # §4.10.x L(n) over read_bool
        x = 0
        for _ in range(n):
            x = 2 * x + self.bits[self.i]
            self.i += 1
        return x

    def NS(self, n):  # §4.10.10
        w = n.bit_length()
        m = (1 << w) - n
        v = self.L(w - 1)
        if v < m:
            return v
        return (v << 1) - m + self.L(1)


def decode_subexp(rd, num_syms, k):  # §5.9.28
    i, mk = 0, 0
    while True:
        b2 = k + i - 1 if i else k
        a = 1 << b2
        if num_syms <= mk + 3 * a:
            return rd.NS(num_syms - mk) + mk
        if rd.L(1):
            i += 1
            mk += a
        else:
            return rd.L(b2) + mk


def inverse_recenter(r, v):  # §5.9.29
    if v > 2 * r:
        return v
    if v & 1:
        return r - ((v + 1) >> 1)
    return r + (v >> 1)


def decode_signed_ref(rd, low, high, k, r):  # §5.9.26-27
    mx, r = high - low, r - low
    v = decode_subexp(rd, mx, k)
    x = inverse_recenter(r, v) if (r << 1) <= mx else mx - 1 - inverse_recenter(mx - 1 - r, v)
    return x + low


@pytest.mark.parametrize("low,high,k", [(-5, 11, 1), (-23, 9, 2), (-17, 47, 3), (-96, 32, 4), (-32, 96, 4)])
def test_subexp_codes_decode_back(oo, low, high, k):
    bits = (C.c_uint8 * 64)()
    cost = C.c_uint64()
    for r in sorted({low, high - 1, (low + high) // 2, low + 1, high - 2}):
        for v in range(low, high):
            n = oo.oo_cw_signed_ref_bits(low, high, k, r, v, bits, 64, C.byref(cost))
            rd = BitReader(bits[:n])
            assert decode_signed_ref(rd, low, high, k, r) == v and rd.i == n, (r, v)
            assert abs(cost.value - 256 * n) <= n + 2


# ---- the election ---------------------------------------------------------

def _unit(t):
    u = abi.LrUnit(type=t[0], sgr_set=t[1])
    for a in range(2):
        for b in range(3):
            u.wiener[a][b] = t[2][a][b]
    u.sgr_xqd[0], u.sgr_xqd[1] = t[3]
    return u


def check_choices(oo, inp, res):
    cdf = (C.c_uint16 * 4)(*inp.restoration_type_cdf)
    d_total = d_pre = 0
    any_lr = []
    for p in range(inp.num_planes):
        fc = 1 if p else 0
        ref_w = [list(WIENER_MID), list(WIENER_MID)]
        ref_x = [-32, 31]
        kinds = set()
        for k, c in enumerate(res.cands[p]):
            d_pre += c["d_none"]
            sbest = min(range(16), key=lambda s: (c["d_sgr"][s], s))
            wt = (abi.RESTORE_WIENER, 0, c["wiener"], (0, 0))
            st = (abi.RESTORE_SGRPROJ, sbest, ((0, 0, 0), (0, 0, 0)), c["sgr_xqd"][sbest])
            nt = (0, 0, ((0, 0, 0), (0, 0, 0)), (0, 0))
            rw = (C.c_int32 * 6)(*(ref_w[0] + ref_w[1]))
            rx = (C.c_int32 * 2)(*ref_x)
            price = [oo.oo_lr_price_unit(rw, rx, p, C.byref(_unit(t)), cdf) for t in (nt, wt, st)]
            ds = [c["d_none"], c["d_wiener"], c["d_sgr"][sbest]]
            scores = [d * 256 + inp.lambda_ * r for d, r in zip(ds, price)]
            best = 0
            for i in (1, 2):
                if scores[i] < scores[best]:
                    best = i
            chosen = (nt, wt, st)[best]
            got = res.units[p][k]
            assert got == chosen, (p, k, got, chosen, scores)
            d_total += ds[best]
            kinds.add(chosen[0])
            # the running references advance with the coded unit
            if chosen[0] == abi.RESTORE_WIENER:
                for pass_ in range(2):
                    for j in range(fc, 3):
                        ref_w[pass_][j] = chosen[2][pass_][j]
            elif chosen[0] == abi.RESTORE_SGRPROJ:
                ref_x = list(chosen[3])
        w, s = abi.RESTORE_WIENER in kinds, abi.RESTORE_SGRPROJ in kinds
        frt = 0 if not (w or s) else (abi.RESTORE_WIENER if not s else (abi.RESTORE_SGRPROJ if not w else 3))
        assert res.frame_restoration_type[p] == frt
        any_lr.append(frt != 0)
    assert res.d == d_total and res.d_pre == d_pre
    assert res.elected == any(any_lr) and res.uses_lr == int(any(any_lr))
    assert res.uses_chroma_lr == int(any(any_lr[1:]))
    assert list(res.loop_restoration_size[:inp.num_planes]) == [64 if res.elected else 0] * inp.num_planes


def test_choices_are_first_argmin(case, oo):
    inp, res = case
    check_choices(oo, inp, res)


def test_lambda_extremes(oo):
    inp = synthetic_lr_input(11, 100, 76, 10, 1, 1, lambda_=0)
    res = elect_lr(inp, oo, "oo_lr_elect")
    check_choices(oo, inp, res)
    for p in range(inp.num_planes):
        for k, c in enumerate(res.cands[p]):
            best = min(c["d_none"], c["d_wiener"], min(c["d_sgr"]))
            t = res.units[p][k][0]
            d = {0: c["d_none"], 1: c["d_wiener"], 2: c["d_sgr"][res.units[p][k][1]]}[t]
            assert d == best
    inp.lambda_ = 1 << 40
    res = elect_lr(inp, oo, "oo_lr_elect")
    assert not res.elected and res.d == res.d_pre
    assert all(u[0] == 0 for p in range(3) for u in res.units[p])


def test_apply_plan_reproduces_d(case, oo):
    inp, res = case
    recon, ssd = apply_lr_plan(res, inp, oo, "oo_lr_apply_plan")
    assert ssd == res.d
    if res.elected:
        grids = [res.raw_units[p] for p in range(inp.num_planes)]
        out = lr_frame(oo, inp, grids, frt=res.frame_restoration_type)
        for p in range(inp.num_planes):
            expect = out[p] if res.frame_restoration_type[p] else inp.cdef[p]
            assert (recon[p] == expect).all()


def test_detail_content_elects_restoration(oo):
    # tests/lr_ab.rs: 192x128 detail content, q = 140 (lambda 1 + q^2/32)
    inp = synthetic_lr_input(5, 192, 128, 8, 1, 1, lambda_=1 + 140 * 140 // 32, blur=1.5, noise=0.5)
    res = elect_lr(inp, oo, "oo_lr_elect")
    assert res.elected and res.uses_lr and res.frame_restoration_type[0] != 0
    assert res.d < res.d_pre
