"""Hand-derived deblocking cases that pin the edge decisions of
loop_filter_edge (loop_filter.rs:302-423) output-exactly.

Each case is a tiny frame whose expected output is built from the
sample-level primitives (filter_mask / narrow_filter / wide_filter, pinned
by their own known-answer tests in test_oracle_kat.py) applied at the edge
positions, filter sizes and strengths derived BY HAND from §7.14.2 for
that geometry.  The frame-level walk must then produce exactly that plane;
every other sample must stay as it was.  A mistake in the edge decisions
(onScreen gates, chroma planeSize, isTxEdge / applyFilter, filter size from
both sides, the prev-side level fallback) changes where or how an edge is
filtered and fails the comparison; test_oracle_lf_edges.py shows each such
mutation of the oracle failing.

The cases are shared by the CPU pins (oracle) and the GPU drop-in parity
tests (tests/test_gpu_lf_edges.py).
"""
import ctypes as C

import numpy as np

from oxav1 import abi
from oxav1.frames import planes_struct

# BLOCK_* and TX_* ordinals (cdf.rs:14723-14770, 531-557)
BLOCK_8X8, BLOCK_16X16, BLOCK_32X8 = 3, 6, 19
TX_4X4, TX_8X8, TX_16X16, TX_16X4, TX_32X8 = 0, 1, 2, 14, 16


class LfCase:
    """One frame: context, dense grids, input planes and the expected
    (plane, pass, x, y, filter_size, (limit, blimit, thresh)) edge samples."""

    def __init__(self, name, mi_rows, mi_cols, *, frame_w=None, frame_h=None, num_planes=1,
                 sx=1, sy=1, levels=(10, 10, 0, 0), sharpness=0, bit_depth=8):
        self.name = name
        c = abi.LfCtx()
        for i in range(4):
            c.loop_filter_level[i] = levels[i]
        c.loop_filter_sharpness = sharpness
        c.num_planes, c.bit_depth = num_planes, bit_depth
        c.subsampling_x, c.subsampling_y = (sx, sy) if num_planes > 1 else (1, 1)
        c.mi_rows, c.mi_cols = mi_rows, mi_cols
        c.frame_width = frame_w if frame_w is not None else mi_cols * 4
        c.frame_height = frame_h if frame_h is not None else mi_rows * 4
        self.ctx = c
        n = mi_rows * mi_cols
        self.grid = {k: np.zeros(n, dtype=np.uint8) for k in
                     ("is_intra", "skip", "ref_frame", "mode", "segment_id", "mi_size")}
        self.delta_lf = np.zeros(n * 4, dtype=np.int8)
        self.tx = []
        for p in range(3):
            psx, psy = (c.subsampling_x, c.subsampling_y) if p else (0, 0)
            self.tx.append(np.zeros(((mi_rows + psy) >> psy) * ((mi_cols + psx) >> psx),
                                    dtype=np.uint8))
        self.seg_active = np.zeros((8, 8), dtype=np.uint8)
        self.seg_data = np.zeros((8, 8), dtype=np.int16)
        self.planes = []
        self.edges = []  # (plane, pass, x, y, filter_size, strength)

    def mi(self, key):
        c = self.ctx
        return self.grid[key].reshape(c.mi_rows, c.mi_cols)

    def tx_grid(self, p):
        c = self.ctx
        psx, psy = (c.subsampling_x, c.subsampling_y) if p else (0, 0)
        return self.tx[p].reshape((c.mi_rows + psy) >> psy, (c.mi_cols + psx) >> psx)

    def grids(self):
        """abi.LfGrids over this case's arrays (they must stay alive)."""
        g = abi.LfGrids()
        for k, v in self.grid.items():
            setattr(g, k, v.ctypes.data_as(C.POINTER(C.c_uint8)))
        g.delta_lf = self.delta_lf.ctypes.data_as(C.POINTER(C.c_int8))
        for p in range(3):
            g.lf_tx_size[p] = self.tx[p].ctypes.data_as(C.POINTER(C.c_uint8))
        for s in range(8):
            for f in range(8):
                g.seg_feature_active[s][f] = int(self.seg_active[s, f])
                g.seg_feature_data[s][f] = int(self.seg_data[s, f])
        return g

    def expected(self, prim):
        """Input planes with the listed edge samples filtered in walk order
        (plane, then pass 0 before pass 1) by the sample-level primitives of
        `prim` (the oracle library)."""
        out = [p.copy() for p in self.planes]
        structs = [planes_struct([a]) for a in out]
        bd = self.ctx.bit_depth
        for plane in range(len(out)):
            for pss in (0, 1):
                for (pl, ps_, x, y, fs, (limit, blimit, thresh)) in self.edges:
                    if pl != plane or ps_ != pss:
                        continue
                    dx, dy = (1, 0) if pss == 0 else (0, 1)
                    ps = structs[plane]
                    # §7.14.6.1 dispatch (loop_filter.rs:582-618) over the pinned primitives
                    m = prim.oo_lf_filter_mask(ps, x, y, plane, limit, blimit, thresh, dx, dy,
                                               fs, bd)
                    if not m.filter_mask:
                        continue
                    if fs == 4 or not m.flat_mask:
                        prim.oo_lf_narrow_filter(ps, x, y, dx, dy, int(m.hev_mask), bd)
                    elif fs == 8 or not m.flat_mask2:
                        prim.oo_lf_wide_filter(ps, x, y, plane, dx, dy, 3)
                    else:
                        prim.oo_lf_wide_filter(ps, x, y, plane, dx, dy, 4)
        return out

    def run(self, fn):
        """Run a loop_filter_frame-shaped entry point on a copy of the planes."""
        out = [p.copy() for p in self.planes]
        g = self.grids()
        ret = fn(C.byref(self.ctx), C.byref(g), planes_struct(out), len(out))
        return ret, out


def strength(lvl, sharpness=0):
    """(limit, blimit, thresh) of §7.14.4 for a level, as the strength known
    answers state them (loop_filter.rs:982-1054)."""
    shift = 2 if sharpness > 4 else (1 if sharpness > 0 else 0)
    lim = max(lvl >> shift, 1)
    if sharpness:
        lim = min(lim, 9 - sharpness)
    return lim, 2 * (lvl + 2) + lim, lvl >> 4


def _edge_run(case, plane, pss, xp, yp, fs, st, n=4):
    """The four samples of one 4-sample edge segment at plane position (xp, yp)."""
    for i in range(n):
        x, y = (xp, yp + i) if pss == 0 else (xp + i, yp)
        case.edges.append((plane, pss, x, y, fs, st))


def case_2x2_grid():
    """loop_filter.rs:1183-1233: a 2x2 mi grid, intra, TX_4X4 in BLOCK_8X8.
    The interior vertical edge (x = 4, rows 0..7) and horizontal edge (y = 4,
    cols 0..7) are tx edges and intra forces applyFilter; edges at x = 0 and
    y = 0 are off by the onScreen gate.  Filter size 4 (narrow)."""
    k = LfCase("2x2_grid", 2, 2, levels=(10, 10, 0, 0))
    k.grid["is_intra"][:] = 1
    k.grid["mi_size"][:] = BLOCK_8X8
    a = np.full((8, 8), 100, dtype=np.int32)
    a[:, 4:] = 104
    a[4:, :] += 2
    k.planes = [a]
    st = strength(10)
    for r in (0, 1):
        _edge_run(k, 0, 0, 4, 4 * r, 4, st)
    for c in (0, 1):
        _edge_run(k, 0, 1, 4 * c, 4, 4, st)
    return k


def case_on_screen_gate():
    """Coded 12x12 inside a 16x16 mi-padded plane (mi 4x4): edges at x = 12
    or y = 12 are off screen (x >= FrameWidth, loop_filter.rs:324-335) and
    stay untouched although the padded samples there hold steps."""
    k = LfCase("on_screen_gate", 4, 4, frame_w=12, frame_h=12, levels=(12, 12, 0, 0))
    k.grid["is_intra"][:] = 1
    k.grid["mi_size"][:] = BLOCK_8X8
    yy, xx = np.mgrid[0:16, 0:16]
    k.planes = [(90 + 3 * (xx // 4) + 2 * (yy // 4)).astype(np.int32)]
    st = strength(12)
    for col in (1, 2):          # x = 4, 8 (x = 12 gated)
        for row in range(4):    # every row of the plane (y < FrameHeight only gates the cell)
            if row * 4 < 12:
                _edge_run(k, 0, 0, 4 * col, 4 * row, 4, st)
    for row in (1, 2):          # y = 4, 8
        for col in range(3):    # x = 0, 4, 8 (x = 12 gated)
            _edge_run(k, 0, 1, 4 * col, 4 * row, 4, st)
    return k


def case_chroma_plane_size():
    """4:2:0, BLOCK_32X8 leaves, skip inter (applyFilter needs isBlockEdge).
    Luma: TX_32X8, vertical block/tx edges every 32 (filter size 16),
    horizontal every 8 (size 8).  Chroma: the plane residual size is
    BLOCK_16X4 (cdf.rs:1250-1275) and the chroma tx TX_16X4, so horizontal
    chroma edges fire every 4 plane rows (size 4) and vertical ones every 16
    (size min(8, 16) = 8).  planeSize = MiSize would make the chroma block
    height 8 and drop the rows at 4 and 12 (the r387 fix,
    loop_filter.rs:356-373)."""
    k = LfCase("chroma_plane_size", 8, 16, num_planes=3, levels=(20, 18, 16, 14))
    k.grid["mi_size"][:] = BLOCK_32X8
    k.grid["skip"][:] = 1
    k.grid["ref_frame"][:] = 1
    k.grid["mode"][:] = 14
    k.tx[0][:] = TX_32X8
    k.tx[1][:] = TX_16X4
    k.tx[2][:] = TX_16X4
    yy, xx = np.mgrid[0:32, 0:64]
    k.planes = [(120 + 2 * (xx // 32) + 3 * (yy // 8) + ((xx * 7 + yy * 3) % 3)).astype(np.int32)]
    cy, cx = np.mgrid[0:16, 0:32]
    k.planes.append((140 + 2 * (cx // 16) + 2 * (cy // 4) + ((cx + cy) % 2)).astype(np.int32))
    k.planes.append((60 + 3 * (cx // 16) + 2 * (cy // 4)).astype(np.int32))
    # luma
    for row in range(8):
        _edge_run(k, 0, 0, 32, 4 * row, 16, strength(20))
    for yp in (8, 16, 24):
        for col in range(16):
            _edge_run(k, 0, 1, 4 * col, yp, 8, strength(18))
    # chroma: levels[2] for Cb, levels[3] for Cr
    for plane, lvl in ((1, 16), (2, 14)):
        for cyp in range(0, 16, 4):
            _edge_run(k, plane, 0, 16, cyp, 8, strength(lvl))
        for yp in (4, 8, 12):
            for cxp in range(0, 32, 4):
                _edge_run(k, plane, 1, cxp, yp, 4, strength(lvl))
    return k


def case_prev_level_fallback():
    """Current side at level 0 (segment 1 with ALT_LF_Y_V = -63), prev side
    at level 10: the edge is filtered with the prev side's strengths
    (loop_filter.rs:391-396).  The second edge has the zero-level block on
    the prev side and filters with the current side's level.  Pass 1 has
    level 0 everywhere and filters nothing."""
    k = LfCase("prev_level_fallback", 2, 4, levels=(10, 0, 0, 0))
    k.grid["is_intra"][:] = 1
    k.grid["mi_size"][:] = BLOCK_8X8
    k.mi("segment_id")[:, 1] = 1      # x = 4..7: level 0
    k.seg_active[1, 1] = 1            # SEG_LVL_ALT_LF_Y_V
    k.seg_data[1, 1] = -63
    yy, xx = np.mgrid[0:8, 0:16]
    k.planes = [(100 + 4 * (xx // 4) + (yy % 2)).astype(np.int32)]
    st = strength(10)
    for row in (0, 1):
        _edge_run(k, 0, 0, 4, 4 * row, 4, st)   # current lvl 0 -> prev (col 0) lvl 10
        _edge_run(k, 0, 0, 8, 4 * row, 4, st)   # prev lvl 0, current (col 2) lvl 10
        _edge_run(k, 0, 0, 12, 4 * row, 4, st)  # both 10
    return k


def case_tx_edges_and_skip(skip, intra):
    """Two BLOCK_16X16 blocks side by side (32x16 plane): the left carries
    TX_8X8, the right TX_16X16.  Pass 0 tx edges: x = 8 (inside the left
    block) and x = 16 (also a block edge); x = 24 is not a tx edge of the
    right block's 16-wide transform.  Pass 1: y = 8 is a tx edge of the left
    block only.  With skip inter blocks only block edges filter
    (applyFilter, loop_filter.rs:377-387); intra or non-skip blocks filter
    every tx edge.  The filter size at x = 16 is min(TxW[TX_16X16],
    TxW[TX_8X8]) = 8, taken from both sides (loop_filter.rs:434-442)."""
    k = LfCase(f"tx_edges_skip{skip}_intra{intra}", 4, 8, levels=(16, 16, 0, 0))
    k.grid["mi_size"][:] = BLOCK_16X16
    k.grid["skip"][:] = skip
    k.grid["is_intra"][:] = intra
    k.grid["ref_frame"][:] = 0 if intra else 2
    k.grid["mode"][:] = 0 if intra else 15
    t = k.tx_grid(0)
    t[:, 0:4] = TX_8X8
    t[:, 4:8] = TX_16X16
    # flat runs either side of each edge (76 | 80 | 100 across x = 8 and 16,
    # +3 below y = 8), so the flat masks hold and a 16-sample filter would
    # differ from the 8-sample one at x = 16
    yy, xx = np.mgrid[0:16, 0:32]
    k.planes = [np.where(xx < 8, 76, np.where(xx < 16, 80, 100)).astype(np.int32) +
                3 * (yy >= 8).astype(np.int32)]
    st = strength(16)
    filt_all = (not skip) or intra
    for row in range(4):
        if filt_all:
            _edge_run(k, 0, 0, 8, 4 * row, 8, st)
        _edge_run(k, 0, 0, 16, 4 * row, 8, st)
    if filt_all:
        for col in range(4):
            _edge_run(k, 0, 1, 4 * col, 8, 8, st)
    return k


def case_chroma_odd_cell():
    """4:2:0 over BLOCK_4X4 leaves: a chroma edge takes its mode info from
    the odd (bottom-right) mi cell of each 2x2 luma group (row |= subY,
    col |= subX, loop_filter.rs:318-319).  Those odd cells carry segment 1
    whose ALT_LF_U feature drops the Cb level to 0 (and the prev side, also
    an odd cell, is 0 too), so Cb stays untouched while Cr filters every
    4-sample tx edge.  Reading the even cells would filter Cb."""
    k = LfCase("chroma_odd_cell", 4, 4, num_planes=3, levels=(0, 0, 20, 20))
    k.grid["is_intra"][:] = 1
    k.grid["mi_size"][:] = 0  # BLOCK_4X4
    k.mi("segment_id")[1::2, 1::2] = 1
    k.seg_active[1, 3] = 1     # SEG_LVL_ALT_LF_U
    k.seg_data[1, 3] = -63
    yy, xx = np.mgrid[0:16, 0:16]
    k.planes = [(100 + (xx + yy) % 5).astype(np.int32)]
    cy, cx = np.mgrid[0:8, 0:8]
    k.planes.append((100 + 3 * (cx // 4) + 2 * (cy // 4)).astype(np.int32))
    k.planes.append((90 + 2 * (cx // 4) + 3 * (cy // 4)).astype(np.int32))
    st = strength(20)
    for yp in (0, 4):
        _edge_run(k, 2, 0, 4, yp, 4, st)
    for xp in (0, 4):
        _edge_run(k, 2, 1, xp, 4, 4, st)
    return k


def all_cases():
    return [case_2x2_grid(), case_on_screen_gate(), case_chroma_plane_size(),
            case_prev_level_fallback(), case_tx_edges_and_skip(1, 0),
            case_tx_edges_and_skip(0, 0), case_tx_edges_and_skip(1, 1), case_chroma_odd_cell()]
