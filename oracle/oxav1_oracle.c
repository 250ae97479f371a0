/*
 * oxav1_oracle.c — CPU oracle: literal C restatement of the reference's
 * filter path.  TEST INFRASTRUCTURE ONLY (see oxav1_oracle.h).
 *
 * Every function cites the reference file:line it restates.  Integer widths
 * follow the reference: i32 everywhere except the i64 sites
 * (loop_filter.rs:811-817 wide-filter accumulator, superres.rs:363-375
 * setup, loop_restoration.rs:615-666 box filter).
 *
 * Parity pins (tests/test_oracle_*.py):
 *   - every known-answer unit test of the five reference modules, ported;
 *   - the spec tables against values parsed from the reference sources;
 *   - closed forms / invariants (identity filters, flat planes, table sums);
 *   - independent restatements on small inputs: the oracle's deblocker run
 *     sequentially vs. all-edges-parallel, CDEF direction search vs. a
 *     brute-force numpy argmax, superres vs. a numpy polyphase convolution,
 *     LR box sums vs. numpy window sums, film-grain LFSR vs. an
 *     independent bit-serial model;
 *   - numpy models written from the spec equations for the Wiener and
 *     self-guided filters on non-uniform input, the chroma grain
 *     recurrence with its luma tap, and the noise stripes / overlap
 *     blending / scaled blend of film-grain synthesis.
 */
#define _POSIX_C_SOURCE 199309L
#include "oxav1_oracle.h"

#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "oo_tables.h"

#define MI_SIZE 4

static inline int32_t clip3(int32_t a, int32_t b, int32_t x) { return x < a ? a : (x > b ? b : x); }
static inline int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }
static inline int32_t imin(int32_t a, int32_t b) { return a < b ? a : b; }
static inline int32_t iabs(int32_t a) { return a < 0 ? -a : a; }
static inline int32_t round2_i32(int32_t x, uint32_t n) {
    return n == 0 ? x : (x + (1 << (n - 1))) >> n;
}
static inline int64_t round2_i64(int64_t x, uint32_t n) {
    return n == 0 ? x : (x + ((int64_t)1 << (n - 1))) >> n;
}
static inline uint32_t round2_u32(uint32_t x, uint32_t n) {
    return n == 0 ? x : (x + (1u << (n - 1))) >> n;
}
static inline uint32_t floor_log2_u32(uint32_t x) {
    return x == 0 ? 0 : 31u - (uint32_t)__builtin_clz(x);
}

/* ===================================================================== */
/* PlaneBuffer get/set (loop_filter.rs:128-158)                           */
/* ===================================================================== */
static inline int32_t pb_get(const oxav1_plane_i32 *b, int32_t y, int32_t x) {
    int32_t yc = clip3(0, (int32_t)b->rows - 1, y);
    int32_t xc = clip3(0, (int32_t)b->cols - 1, x);
    return b->samples[(size_t)yc * b->cols + (size_t)xc];
}
static inline void pb_set(oxav1_plane_i32 *b, int32_t y, int32_t x, int32_t v) {
    if (y < 0 || x < 0) return;
    if ((uint32_t)y >= b->rows || (uint32_t)x >= b->cols) return;
    b->samples[(size_t)y * b->cols + (size_t)x] = v;
}

/* Deblocking reads go through lf_get.  oo_loop_filter_frame_parallel points
 * s_lf_read at a snapshot of the plane taken at the start of each pass, so
 * every edge of the pass reads the pass-start samples (all edges "at once"),
 * while writes still land in the live plane.  Test infrastructure for the
 * sequential-vs-parallel equivalence of the in-place walk. */
static _Thread_local const oxav1_plane_i32 *s_lf_read = NULL;
static inline int32_t lf_get(const oxav1_plane_i32 *b, int32_t y, int32_t x) {
    return pb_get(s_lf_read ? s_lf_read : b, y, x);
}

/* ===================================================================== */
/* cdf.rs helpers                                                        */
/* ===================================================================== */

/* get_plane_residual_size (cdf.rs:1250-1275); -1 = None. */
int32_t oo_get_plane_residual_size(uint32_t subsize, uint8_t plane, uint8_t sx, uint8_t sy) {
    if (subsize >= 22) return -1;
    uint32_t subx = plane > 0 ? (sx != 0) : 0;
    uint32_t suby = plane > 0 ? (sy != 0) : 0;
    uint32_t raw = oo_subsampled_size[subsize * 4 + subx * 2 + suby];
    return raw == OXAV1_BLOCK_INVALID ? -1 : (int32_t)raw;
}

/* get_tx_size (cdf.rs:12976-13012); -1 = None. */
int32_t oo_get_tx_size(uint8_t plane, uint32_t tx_size, uint32_t mi_size, uint8_t sx, uint8_t sy) {
    if (plane > 2 || tx_size >= 19 || mi_size >= 22 || sx > 1 || sy > 1) return -1;
    if (plane == 0) return (int32_t)tx_size;
    int32_t ps = oo_get_plane_residual_size(mi_size, plane, sx, sy);
    if (ps < 0) return -1;
    uint32_t uv = oo_max_tx_size_rect[ps];
    uint32_t w = oo_tx_width[uv], h = oo_tx_height[uv];
    if (w == 64 || h == 64) {
        if (w == 16) return 9;  /* TX_16X32 */
        if (h == 16) return 10; /* TX_32X16 */
        return 3;               /* TX_32X32 */
    }
    return (int32_t)uv;
}

/* ===================================================================== */
/* §7.14 deblocking                                                       */
/* ===================================================================== */

/* Closure accessors over the dense grids (the bridge, cdf.rs:20449-20540). */
static inline size_t mi_idx(const oxav1_lf_ctx *c, uint32_t r, uint32_t col) {
    return (size_t)r * c->mi_cols + col;
}
static uint8_t g_segment(const oxav1_lf_ctx *c, const oxav1_lf_grids *g, uint32_t r, uint32_t col) {
    return g->segment_id ? g->segment_id[mi_idx(c, r, col)] : 255;
}
static int8_t g_delta_lf(const oxav1_lf_ctx *c, const oxav1_lf_grids *g, uint32_t r, uint32_t col,
                         uint32_t idx) {
    return g->delta_lf ? g->delta_lf[mi_idx(c, r, col) * 4 + idx] : 0;
}
static int g_seg_active(const oxav1_lf_grids *g, uint32_t seg, uint32_t feat) {
    if (seg >= 8 || feat >= 8) return 0;
    return g->seg_feature_active[seg][feat] != 0;
}
static int32_t g_seg_data(const oxav1_lf_grids *g, uint32_t seg, uint32_t feat) {
    if (seg >= 8 || feat >= 8) return 0;
    return g->seg_feature_data[seg][feat];
}
static uint32_t g_tx(const oxav1_lf_ctx *c, const oxav1_lf_grids *g, uint32_t plane, uint32_t sr,
                     uint32_t sc) {
    uint32_t sx = plane ? c->subsampling_x : 0;
    uint32_t sy = plane ? c->subsampling_y : 0;
    uint32_t cols = (c->mi_cols + sx) >> sx;
    (void)sy;
    return g->lf_tx_size[plane][(size_t)sr * cols + sc];
}

/* filter_size (loop_filter.rs:434-442). */
uint32_t oo_lf_filter_size(uint32_t tx_sz, uint32_t prev_tx_sz, uint8_t pass, uint8_t plane) {
    uint32_t base = pass == 0 ? (uint32_t)imin(oo_tx_width[tx_sz], oo_tx_width[prev_tx_sz])
                              : (uint32_t)imin(oo_tx_height[tx_sz], oo_tx_height[prev_tx_sz]);
    uint32_t cap = plane == 0 ? 16 : 8;
    return base < cap ? base : cap;
}

/* adaptive_filter_strength_selection (loop_filter.rs:527-572). */
static uint8_t lf_strength_selection(const oxav1_lf_ctx *ctx, const oxav1_lf_grids *g,
                                     uint32_t segment, uint8_t ref_frame, uint8_t mode_type,
                                     int8_t delta_lf, uint8_t plane, uint8_t pass) {
    uint32_t i = plane == 0 ? pass : (uint32_t)(plane + 1);
    int32_t base = clip3(0, 63, (int32_t)delta_lf + (int32_t)ctx->loop_filter_level[i]);
    int32_t lvl_seg = base;
    uint32_t feature = 1 + i;
    if (g_seg_active(g, segment, feature))
        lvl_seg = clip3(0, 63, g_seg_data(g, segment, feature) + lvl_seg);
    if (ctx->loop_filter_delta_enabled) {
        int32_t n_shift = lvl_seg >> 5;
        if (ref_frame == 0) {
            lvl_seg += (int32_t)ctx->loop_filter_ref_deltas[0] * (1 << n_shift);
        } else {
            uint32_t ref_idx = ref_frame < 7 ? ref_frame : 7;
            uint32_t mode_idx = mode_type < 1 ? mode_type : 1;
            lvl_seg += (int32_t)ctx->loop_filter_ref_deltas[ref_idx] * (1 << n_shift);
            lvl_seg += (int32_t)ctx->loop_filter_mode_deltas[mode_idx] * (1 << n_shift);
        }
        lvl_seg = clip3(0, 63, lvl_seg);
    }
    return (uint8_t)clip3(0, 255, lvl_seg);
}

/* adaptive_filter_strength (loop_filter.rs:465-524). */
oo_lf_strength oo_lf_adaptive_filter_strength(const oxav1_lf_ctx *ctx, const oxav1_lf_grids *g,
                                              uint32_t row, uint32_t col, uint8_t plane,
                                              uint8_t pass) {
    size_t k = mi_idx(ctx, row, col);
    uint8_t segment = g_segment(ctx, g, row, col);
    uint8_t r = g->ref_frame[k];
    uint8_t m = g->mode[k];
    uint8_t mode_type = (m >= 14 && m != 16 && m != 24) ? 1 : 0;
    uint32_t dli = ctx->delta_lf_multi ? (plane == 0 ? pass : (uint32_t)(plane + 1)) : 0;
    int8_t dlf = g_delta_lf(ctx, g, row, col, dli);
    uint8_t lvl = lf_strength_selection(ctx, g, segment, r, mode_type, dlf, plane, pass);
    uint8_t sharp = ctx->loop_filter_sharpness;
    int32_t shift = sharp > 4 ? 2 : (sharp > 0 ? 1 : 0);
    int32_t limit_i = sharp > 0 ? clip3(1, 9 - (int32_t)sharp, (int32_t)lvl >> shift)
                                : imax((int32_t)lvl >> shift, 1);
    oo_lf_strength s;
    s.lvl = lvl;
    s.limit = (uint8_t)limit_i;
    s.blimit = (uint8_t)(2 * ((uint32_t)lvl + 2) + (uint32_t)s.limit);
    s.thresh = lvl >> 4;
    return s;
}

static inline int32_t filter4_clamp(int32_t v, uint8_t bd) {
    return clip3(-(1 << (bd - 1)), (1 << (bd - 1)) - 1, v);
}

/* filter_mask (loop_filter.rs:643-736). */
oo_lf_masks oo_lf_filter_mask(const oxav1_plane_i32 *b, int32_t x, int32_t y, uint8_t plane,
                              uint8_t limit, uint8_t blimit, uint8_t thresh, int32_t dx,
                              int32_t dy, uint32_t filter_size, uint8_t bit_depth) {
    int32_t q0 = lf_get(b, y, x), q1 = lf_get(b, y + dy, x + dx);
    int32_t q2 = lf_get(b, y + 2 * dy, x + 2 * dx), q3 = lf_get(b, y + 3 * dy, x + 3 * dx);
    int32_t q4 = lf_get(b, y + 4 * dy, x + 4 * dx), q5 = lf_get(b, y + 5 * dy, x + 5 * dx);
    int32_t q6 = lf_get(b, y + 6 * dy, x + 6 * dx);
    int32_t p0 = lf_get(b, y - dy, x - dx), p1 = lf_get(b, y - 2 * dy, x - 2 * dx);
    int32_t p2 = lf_get(b, y - 3 * dy, x - 3 * dx), p3 = lf_get(b, y - 4 * dy, x - 4 * dx);
    int32_t p4 = lf_get(b, y - 5 * dy, x - 5 * dx), p5 = lf_get(b, y - 6 * dy, x - 6 * dx);
    int32_t p6 = lf_get(b, y - 7 * dy, x - 7 * dx);
    uint32_t bds = (uint32_t)bit_depth - 8;
    int32_t thresh_bd = (int32_t)thresh << bds;
    oo_lf_masks m;
    m.hev_mask = iabs(p1 - p0) > thresh_bd || iabs(q1 - q0) > thresh_bd;
    uint32_t filter_len = filter_size == 4 ? 4 : (plane != 0 ? 6 : (filter_size == 8 ? 8 : 16));
    int32_t limit_bd = (int32_t)limit << bds, blimit_bd = (int32_t)blimit << bds;
    int mask = iabs(p1 - p0) > limit_bd;
    mask |= iabs(q1 - q0) > limit_bd;
    mask |= iabs(p0 - q0) * 2 + iabs(p1 - q1) / 2 > blimit_bd;
    if (filter_len >= 6) {
        mask |= iabs(p2 - p1) > limit_bd;
        mask |= iabs(q2 - q1) > limit_bd;
    }
    if (filter_len >= 8) {
        mask |= iabs(p3 - p2) > limit_bd;
        mask |= iabs(q3 - q2) > limit_bd;
    }
    m.filter_mask = !mask;
    int32_t thr = 1 << bds;
    if (filter_size >= 8) {
        int f = iabs(p1 - p0) > thr;
        f |= iabs(q1 - q0) > thr;
        f |= iabs(p2 - p0) > thr;
        f |= iabs(q2 - q0) > thr;
        if (filter_len >= 8) {
            f |= iabs(p3 - p0) > thr;
            f |= iabs(q3 - q0) > thr;
        }
        m.flat_mask = !f;
    } else {
        m.flat_mask = 0;
    }
    if (filter_size >= 16) {
        int f = iabs(p6 - p0) > thr;
        f |= iabs(q6 - q0) > thr;
        f |= iabs(p5 - p0) > thr;
        f |= iabs(q5 - q0) > thr;
        f |= iabs(p4 - p0) > thr;
        f |= iabs(q4 - q0) > thr;
        m.flat_mask2 = !f;
    } else {
        m.flat_mask2 = 0;
    }
    return m;
}

/* narrow_filter (loop_filter.rs:743-782). */
void oo_lf_narrow_filter(oxav1_plane_i32 *b, int32_t x, int32_t y, int32_t dx, int32_t dy,
                         int hev_mask, uint8_t bd) {
    int32_t half = 0x80 << (bd - 8);
    int32_t q0 = lf_get(b, y, x), q1 = lf_get(b, y + dy, x + dx);
    int32_t p0 = lf_get(b, y - dy, x - dx), p1 = lf_get(b, y - 2 * dy, x - 2 * dx);
    int32_t ps1 = p1 - half, ps0 = p0 - half, qs0 = q0 - half, qs1 = q1 - half;
    int32_t filter = hev_mask ? filter4_clamp(ps1 - qs1, bd) : 0;
    filter = filter4_clamp(filter + 3 * (qs0 - ps0), bd);
    int32_t filter1 = filter4_clamp(filter + 4, bd) >> 3;
    int32_t filter2 = filter4_clamp(filter + 3, bd) >> 3;
    int32_t oq0 = filter4_clamp(qs0 - filter1, bd) + half;
    int32_t op0 = filter4_clamp(ps0 + filter2, bd) + half;
    pb_set(b, y, x, oq0);
    pb_set(b, y - dy, x - dx, op0);
    if (!hev_mask) {
        int32_t f = round2_i32(filter1, 1);
        int32_t oq1 = filter4_clamp(qs1 - f, bd) + half;
        int32_t op1 = filter4_clamp(ps1 + f, bd) + half;
        pb_set(b, y + dy, x + dx, oq1);
        pb_set(b, y - 2 * dy, x - 2 * dx, op1);
    }
}

/* wide_filter (loop_filter.rs:789-822). */
void oo_lf_wide_filter(oxav1_plane_i32 *b, int32_t x, int32_t y, uint8_t plane, int32_t dx,
                       int32_t dy, uint32_t log2_size) {
    int32_t n = log2_size == 4 ? 6 : (plane == 0 ? 3 : 2);
    int32_t n2 = (log2_size == 3 && plane == 0) ? 0 : 1;
    int32_t f[12];
    int idx = 0;
    for (int32_t i = -n; i < n; i++, idx++) {
        int64_t t = 0;
        for (int32_t j = -n; j <= n; j++) {
            int32_t p = clip3(-(n + 1), n, i + j);
            int64_t tap = iabs(j) <= n2 ? 2 : 1;
            t += (int64_t)lf_get(b, y + p * dy, x + p * dx) * tap;
        }
        f[idx] = (int32_t)round2_i64(t, log2_size);
    }
    idx = 0;
    for (int32_t i = -n; i < n; i++, idx++) pb_set(b, y + i * dy, x + i * dx, f[idx]);
}

/* sample_filtering (loop_filter.rs:582-616). */
static void lf_sample_filtering(oxav1_plane_i32 *b, int32_t x, int32_t y, uint8_t plane,
                                uint8_t limit, uint8_t blimit, uint8_t thresh, int32_t dx,
                                int32_t dy, uint32_t filter_size, uint8_t bd) {
    oo_lf_masks m = oo_lf_filter_mask(b, x, y, plane, limit, blimit, thresh, dx, dy, filter_size, bd);
    if (!m.filter_mask) return;
    if (filter_size == 4 || !m.flat_mask)
        oo_lf_narrow_filter(b, x, y, dx, dy, m.hev_mask, bd);
    else if (filter_size == 8 || !m.flat_mask2)
        oo_lf_wide_filter(b, x, y, plane, dx, dy, 3);
    else
        oo_lf_wide_filter(b, x, y, plane, dx, dy, 4);
}

static uint32_t lf_block_width(int32_t ps) {
    if (ps < 0 || ps >= 22) return 4;
    return (uint32_t)oo_num4x4_w[ps] * 4;
}
static uint32_t lf_block_height(int32_t ps) {
    if (ps < 0 || ps >= 22) return 4;
    return (uint32_t)oo_num4x4_h[ps] * 4;
}

/* loop_filter_edge (loop_filter.rs:302-423). */
void oo_loop_filter_edge(const oxav1_lf_ctx *ctx, const oxav1_lf_grids *g,
                         oxav1_plane_i32 *planes, uint32_t n, uint8_t plane, uint8_t pass,
                         uint32_t row, uint32_t col) {
    uint8_t sx = plane ? ctx->subsampling_x : 0, sy = plane ? ctx->subsampling_y : 0;
    int32_t dx = pass == 0 ? 1 : 0, dy = pass == 0 ? 0 : 1;
    uint32_t x = col * MI_SIZE, y = row * MI_SIZE;
    row |= sy;
    col |= sx;
    if (x >= ctx->frame_width) return;
    if (y >= ctx->frame_height) return;
    if (pass == 0 && x == 0) return;
    if (pass == 1 && y == 0) return;
    int32_t xp = (int32_t)(x >> sx), yp = (int32_t)(y >> sy);
    int32_t prev_row = (int32_t)row - (dy << sy);
    int32_t prev_col = (int32_t)col - (dx << sx);
    if (prev_row < 0 || prev_col < 0) return;
    if (row >= ctx->mi_rows || col >= ctx->mi_cols) return;
    uint32_t mi_size = g->mi_size[mi_idx(ctx, row, col)];
    uint32_t tx_sz = g_tx(ctx, g, plane, row >> sy, col >> sx);
    int32_t ps = oo_get_plane_residual_size(mi_size, plane, sx, sy);
    int32_t plane_size = ps >= 0 ? ps : (int32_t)mi_size;
    int skip = g->skip[mi_idx(ctx, row, col)] != 0;
    int is_intra = g->is_intra[mi_idx(ctx, row, col)] != 0;
    uint32_t prev_tx_sz = g_tx(ctx, g, plane, (uint32_t)prev_row >> sy, (uint32_t)prev_col >> sx);
    int is_block_edge, is_tx_edge;
    if (pass == 0) {
        uint32_t bw = lf_block_width(plane_size);
        is_block_edge = ((uint32_t)xp % (bw ? bw : 1)) == 0;
        uint32_t tw = oo_tx_width[tx_sz];
        is_tx_edge = ((uint32_t)xp % (tw ? tw : 1)) == 0;
    } else {
        uint32_t bh = lf_block_height(plane_size);
        is_block_edge = ((uint32_t)yp % (bh ? bh : 1)) == 0;
        uint32_t th = oo_tx_height[tx_sz];
        is_tx_edge = ((uint32_t)yp % (th ? th : 1)) == 0;
    }
    int apply = is_tx_edge ? (is_block_edge || !skip || is_intra) : 0;
    uint32_t fs = oo_lf_filter_size(tx_sz, prev_tx_sz, pass, plane);
    oo_lf_strength s = oo_lf_adaptive_filter_strength(ctx, g, row, col, plane, pass);
    if (s.lvl == 0)
        s = oo_lf_adaptive_filter_strength(ctx, g, (uint32_t)prev_row, (uint32_t)prev_col, plane, pass);
    if (!apply || s.lvl == 0) return;
    if (plane >= n) return;
    oxav1_plane_i32 *pb = &planes[plane];
    for (int32_t i = 0; i < MI_SIZE; i++) {
        int32_t sxp = xp + dy * i, syp = yp + dx * i;
        lf_sample_filtering(pb, sxp, syp, plane, s.limit, s.blimit, s.thresh, dx, dy, fs,
                            ctx->bit_depth);
    }
}

/* loop_filter_frame (loop_filter.rs:271-294). */
void oo_loop_filter_frame(const oxav1_lf_ctx *ctx, const oxav1_lf_grids *g,
                          oxav1_plane_i32 *planes, uint32_t n) {
    uint32_t np = ctx->num_planes < n ? ctx->num_planes : n;
    for (uint32_t plane = 0; plane < np; plane++) {
        if (plane > 0 && ctx->loop_filter_level[1 + plane] == 0) continue;
        uint32_t sx = plane ? ctx->subsampling_x : 0, sy = plane ? ctx->subsampling_y : 0;
        uint32_t row_step = plane == 0 ? 1 : (1u << sy);
        uint32_t col_step = plane == 0 ? 1 : (1u << sx);
        for (uint8_t pass = 0; pass < 2; pass++)
            for (uint32_t row = 0; row < ctx->mi_rows; row += row_step)
                for (uint32_t col = 0; col < ctx->mi_cols; col += col_step)
                    oo_loop_filter_edge(ctx, g, planes, n, (uint8_t)plane, pass, row, col);
    }
}

/* The walk of oo_loop_filter_frame with every edge of a (plane, pass)
 * reading the samples as they were when the pass began. */
void oo_loop_filter_frame_parallel(const oxav1_lf_ctx *ctx, const oxav1_lf_grids *g,
                                   oxav1_plane_i32 *planes, uint32_t n) {
    uint32_t np = ctx->num_planes < n ? ctx->num_planes : n;
    for (uint32_t plane = 0; plane < np; plane++) {
        if (plane > 0 && ctx->loop_filter_level[1 + plane] == 0) continue;
        uint32_t sx = plane ? ctx->subsampling_x : 0, sy = plane ? ctx->subsampling_y : 0;
        uint32_t row_step = plane == 0 ? 1 : (1u << sy);
        uint32_t col_step = plane == 0 ? 1 : (1u << sx);
        size_t cnt = (size_t)planes[plane].rows * planes[plane].cols;
        oxav1_plane_i32 snap = planes[plane];
        snap.samples = (int32_t *)malloc(cnt * sizeof(int32_t) + 4);
        for (uint8_t pass = 0; pass < 2; pass++) {
            memcpy(snap.samples, planes[plane].samples, cnt * sizeof(int32_t));
            s_lf_read = &snap;
            for (uint32_t row = 0; row < ctx->mi_rows; row += row_step)
                for (uint32_t col = 0; col < ctx->mi_cols; col += col_step)
                    oo_loop_filter_edge(ctx, g, planes, n, (uint8_t)plane, pass, row, col);
            s_lf_read = NULL;
        }
        free(snap.samples);
    }
}

/* The walk of oo_loop_filter_frame with the mi rows and columns of each
 * (plane, pass) visited in reverse order: agrees with the forward walk
 * exactly when no edge's filtering depends on the order of the walk. */
void oo_loop_filter_frame_reversed(const oxav1_lf_ctx *ctx, const oxav1_lf_grids *g,
                                   oxav1_plane_i32 *planes, uint32_t n) {
    uint32_t np = ctx->num_planes < n ? ctx->num_planes : n;
    for (uint32_t plane = 0; plane < np; plane++) {
        if (plane > 0 && ctx->loop_filter_level[1 + plane] == 0) continue;
        uint32_t sx = plane ? ctx->subsampling_x : 0, sy = plane ? ctx->subsampling_y : 0;
        uint32_t row_step = plane == 0 ? 1 : (1u << sy);
        uint32_t col_step = plane == 0 ? 1 : (1u << sx);
        uint32_t nr = (ctx->mi_rows + row_step - 1) / row_step, nc = (ctx->mi_cols + col_step - 1) / col_step;
        for (uint8_t pass = 0; pass < 2; pass++)
            for (uint32_t ri = nr; ri-- > 0;)
                for (uint32_t ci = nc; ci-- > 0;)
                    oo_loop_filter_edge(ctx, g, planes, n, (uint8_t)plane, pass, ri * row_step,
                                        ci * col_step);
    }
}

/* ===================================================================== */
/* §7.15 CDEF                                                             */
/* ===================================================================== */

static const int32_t CDEF_PRI_TAPS[2][2] = {{4, 2}, {3, 3}}; /* cdef.rs:66 */
static const int32_t CDEF_SEC_TAPS[2][2] = {{2, 1}, {2, 1}}; /* cdef.rs:71 */

static int cdef_skip_at(const oxav1_cdef_ctx *c, const oxav1_cdef_grids *g, uint32_t r, uint32_t col) {
    return g->skip[(size_t)r * c->mi_cols + col] != 0;
}
static int8_t cdef_idx_at(const oxav1_cdef_ctx *c, const oxav1_cdef_grids *g, uint32_t r, uint32_t col) {
    if (r >= c->mi_rows || col >= c->mi_cols) return -1;
    uint32_t ucols = (c->mi_cols + 15) >> 4;
    return g->cdef_idx[(size_t)(r >> 4) * ucols + (col >> 4)];
}

/* constrain (cdef.rs:592-606). */
int32_t oo_cdef_constrain(int32_t diff, int32_t threshold, int32_t damping) {
    if (threshold == 0) return 0;
    int32_t adj = imax(0, damping - (int32_t)floor_log2_u32((uint32_t)threshold));
    int32_t sign = diff < 0 ? -1 : 1;
    int32_t ad = iabs(diff);
    int32_t shifted = adj >= 31 ? 0 : ad >> adj;
    int32_t inner = imax(threshold - shifted, 0);
    return sign * clip3(0, ad, inner);
}

/* cdef_copy_plane (cdef.rs:320-352). */
static void cdef_copy_plane(const oxav1_plane_i32 *src, uint32_t nsrc, oxav1_plane_i32 *dst,
                            uint32_t ndst, uint32_t p, uint32_t r, uint32_t c, uint32_t sx,
                            uint32_t sy) {
    if (p >= nsrc || p >= ndst) return;
    const oxav1_plane_i32 *s = &src[p];
    oxav1_plane_i32 *d = &dst[p];
    int32_t y0 = (int32_t)((r * MI_SIZE) >> sy), y1 = (int32_t)(((r + 2) * MI_SIZE) >> sy);
    int32_t x0 = (int32_t)((c * MI_SIZE) >> sx), x1 = (int32_t)(((c + 2) * MI_SIZE) >> sx);
    int32_t cols = (int32_t)(s->cols < d->cols ? s->cols : d->cols);
    int32_t rows = (int32_t)(s->rows < d->rows ? s->rows : d->rows);
    for (int32_t y = y0; y < y1; y++) {
        if (y < 0 || y >= rows) continue;
        for (int32_t x = x0; x < x1; x++) {
            if (x < 0 || x >= cols) continue;
            d->samples[(size_t)y * d->cols + x] = s->samples[(size_t)y * s->cols + x];
        }
    }
}

/* cdef_direction (cdef.rs:366-449). */
void oo_cdef_direction(const oxav1_cdef_ctx *ctx, const oxav1_plane_i32 *src, uint32_t nsrc,
                       uint32_t r, uint32_t c, int32_t *y_dir, int32_t *var) {
    if (nsrc == 0) {
        *y_dir = 0;
        *var = 0;
        return;
    }
    const oxav1_plane_i32 *p0 = &src[0];
    int32_t partial[8][15];
    int32_t cost[8];
    memset(partial, 0, sizeof partial);
    memset(cost, 0, sizeof cost);
    int32_t x0 = (int32_t)c << 2, y0 = (int32_t)r << 2;
    int32_t shift = (int32_t)ctx->bit_depth - 8;
    for (int32_t i = 0; i < 8; i++) {
        for (int32_t j = 0; j < 8; j++) {
            int32_t sx = clip3(0, (int32_t)p0->cols - 1, x0 + j);
            int32_t sy = clip3(0, (int32_t)p0->rows - 1, y0 + i);
            int32_t raw = p0->samples[(size_t)sy * p0->cols + sx];
            int32_t x = (raw >> shift) - 128;
            partial[0][i + j] += x;
            partial[1][i + j / 2] += x;
            partial[2][i] += x;
            partial[3][3 + i - j / 2] += x;
            partial[4][7 + i - j] += x;
            partial[5][3 - i / 2 + j] += x;
            partial[6][j] += x;
            partial[7][i / 2 + j] += x;
        }
    }
    for (int k = 0; k < 8; k++) {
        cost[2] += partial[2][k] * partial[2][k];
        cost[6] += partial[6][k] * partial[6][k];
    }
    cost[2] *= oo_div_table[8];
    cost[6] *= oo_div_table[8];
    for (int i = 0; i < 7; i++) {
        cost[0] += (partial[0][i] * partial[0][i] + partial[0][14 - i] * partial[0][14 - i]) *
                   oo_div_table[i + 1];
        cost[4] += (partial[4][i] * partial[4][i] + partial[4][14 - i] * partial[4][14 - i]) *
                   oo_div_table[i + 1];
    }
    cost[0] += partial[0][7] * partial[0][7] * oo_div_table[8];
    cost[4] += partial[4][7] * partial[4][7] * oo_div_table[8];
    for (int i = 1; i < 8; i += 2) {
        for (int j = 0; j < 5; j++) cost[i] += partial[i][3 + j] * partial[i][3 + j];
        cost[i] *= oo_div_table[8];
        for (int j = 0; j < 3; j++)
            cost[i] += (partial[i][j] * partial[i][j] + partial[i][10 - j] * partial[i][10 - j]) *
                       oo_div_table[2 * j + 2];
    }
    int32_t best = 0, dir = 0;
    for (int k = 0; k < 8; k++) {
        if (cost[k] > best) {
            best = cost[k];
            dir = k;
        }
    }
    *y_dir = dir;
    *var = (best - cost[(dir + 4) & 7]) >> 10;
}

/* cdef_sample_at (cdef.rs:564-585); returns 0 when unavailable. */
static int cdef_sample_at(const oxav1_plane_i32 *s, int32_t pw, int32_t ph, int32_t x0, int32_t y0,
                          int32_t i, int32_t j, int32_t dir, int k, int32_t sign, int32_t *out) {
    int32_t dy = oo_cdef_directions[((dir & 7) * 2 + k) * 2 + 0];
    int32_t dx = oo_cdef_directions[((dir & 7) * 2 + k) * 2 + 1];
    int32_t y = y0 + i + sign * dy, x = x0 + j + sign * dx;
    if (y < 0 || y >= ph || x < 0 || x >= pw) return 0;
    *out = s->samples[(size_t)y * s->cols + x];
    return 1;
}

/* cdef_filter_block (cdef.rs:465-560). */
void oo_cdef_filter_block(const oxav1_cdef_ctx *ctx, const oxav1_plane_i32 *src, uint32_t nsrc,
                          oxav1_plane_i32 *dst, uint32_t ndst, uint8_t plane, uint32_t r,
                          uint32_t c, int32_t pri_str, int32_t sec_str, int32_t damping,
                          int32_t dir) {
    if (plane >= nsrc || plane >= ndst) return;
    const oxav1_plane_i32 *s = &src[plane];
    oxav1_plane_i32 *d = &dst[plane];
    uint32_t sx = plane ? ctx->subsampling_x : 0, sy = plane ? ctx->subsampling_y : 0;
    int32_t coeff_shift = (int32_t)ctx->bit_depth - 8;
    int32_t x0 = ((int32_t)c * MI_SIZE) >> sx, y0 = ((int32_t)r * MI_SIZE) >> sy;
    int32_t w = 8 >> sx, h = 8 >> sy;
    int pri_row = (pri_str >> coeff_shift) & 1;
    int sec_row = pri_row;
    int32_t pw = (int32_t)s->cols, ph = (int32_t)s->rows;
    for (int32_t i = 0; i < h; i++) {
        for (int32_t j = 0; j < w; j++) {
            int32_t py = y0 + i, px = x0 + j;
            if (py < 0 || py >= ph || px < 0 || px >= pw) continue;
            int32_t x = s->samples[(size_t)py * pw + px];
            int32_t sum = 0, mx = x, mn = x, v;
            for (int k = 0; k < 2; k++) {
                for (int si = 0; si < 2; si++) {
                    int32_t sign = si == 0 ? -1 : 1;
                    if (cdef_sample_at(s, pw, ph, x0, y0, i, j, dir, k, sign, &v)) {
                        sum += CDEF_PRI_TAPS[pri_row][k] * oo_cdef_constrain(v - x, pri_str, damping);
                        if (v > mx) mx = v;
                        if (v < mn) mn = v;
                    }
                    for (int oi = 0; oi < 2; oi++) {
                        int32_t dir2 = (dir + (oi == 0 ? -2 : 2)) & 7;
                        if (cdef_sample_at(s, pw, ph, x0, y0, i, j, dir2, k, sign, &v)) {
                            sum += CDEF_SEC_TAPS[sec_row][k] * oo_cdef_constrain(v - x, sec_str, damping);
                            if (v > mx) mx = v;
                            if (v < mn) mn = v;
                        }
                    }
                }
            }
            int32_t delta = (8 + sum - (sum < 0)) >> 4;
            d->samples[(size_t)py * d->cols + px] = clip3(mn, mx, x + delta);
        }
    }
}

/* cdef_block (cdef.rs:220-314). */
static void cdef_block(const oxav1_cdef_ctx *ctx, const oxav1_cdef_grids *g,
                       const oxav1_plane_i32 *src, uint32_t nsrc, oxav1_plane_i32 *dst,
                       uint32_t ndst, uint32_t np, uint32_t r, uint32_t c, int8_t idx) {
    cdef_copy_plane(src, nsrc, dst, ndst, 0, r, c, 0, 0);
    if (np > 1) {
        cdef_copy_plane(src, nsrc, dst, ndst, 1, r, c, ctx->subsampling_x, ctx->subsampling_y);
        cdef_copy_plane(src, nsrc, dst, ndst, 2, r, c, ctx->subsampling_x, ctx->subsampling_y);
    }
    if (idx < 0) return;
    int32_t coeff_shift = (int32_t)ctx->bit_depth - 8;
    int skip = cdef_skip_at(ctx, g, r, c) &&
               (r + 1 >= ctx->mi_rows || cdef_skip_at(ctx, g, r + 1, c)) &&
               (c + 1 >= ctx->mi_cols || cdef_skip_at(ctx, g, r, c + 1)) &&
               (r + 1 >= ctx->mi_rows || c + 1 >= ctx->mi_cols || cdef_skip_at(ctx, g, r + 1, c + 1));
    if (skip) return;
    int32_t y_dir, var;
    oo_cdef_direction(ctx, src, nsrc, r, c, &y_dir, &var);
    uint32_t ui = (uint32_t)idx;
    int32_t pri_y = (int32_t)ctx->params.y_pri_strength[ui] << coeff_shift;
    int32_t sec_y = (int32_t)ctx->params.y_sec_strength[ui] << coeff_shift;
    int32_t dir_y = pri_y == 0 ? 0 : y_dir;
    int32_t var_str = (var >> 6) > 0 ? imin((int32_t)floor_log2_u32((uint32_t)(var >> 6)), 12) : 0;
    int32_t pri_y_scaled = var != 0 ? (pri_y * (4 + var_str) + 8) >> 4 : 0;
    int32_t damping_y = (int32_t)ctx->params.cdef_damping + coeff_shift;
    oo_cdef_filter_block(ctx, src, nsrc, dst, ndst, 0, r, c, pri_y_scaled, sec_y, damping_y, dir_y);
    if (np == 1) return;
    int32_t pri_uv = (int32_t)ctx->params.uv_pri_strength[ui] << coeff_shift;
    int32_t sec_uv = (int32_t)ctx->params.uv_sec_strength[ui] << coeff_shift;
    uint32_t ssx = ctx->subsampling_x ? 1 : 0, ssy = ctx->subsampling_y ? 1 : 0;
    int32_t dir_uv = pri_uv == 0 ? 0 : (int32_t)oo_cdef_uv_dir[(ssx * 2 + ssy) * 8 + (uint32_t)y_dir];
    int32_t damping_uv = (int32_t)ctx->params.cdef_damping + coeff_shift - 1;
    oo_cdef_filter_block(ctx, src, nsrc, dst, ndst, 1, r, c, pri_uv, sec_uv, damping_uv, dir_uv);
    oo_cdef_filter_block(ctx, src, nsrc, dst, ndst, 2, r, c, pri_uv, sec_uv, damping_uv, dir_uv);
}

/* cdef_frame (cdef.rs:180-201). */
void oo_cdef_frame(const oxav1_cdef_ctx *ctx, const oxav1_cdef_grids *g,
                   const oxav1_plane_i32 *src, uint32_t nsrc, oxav1_plane_i32 *dst,
                   uint32_t ndst) {
    uint32_t np = ctx->num_planes;
    if (np > nsrc) np = nsrc;
    if (np > ndst) np = ndst;
    for (uint32_t r = 0; r < ctx->mi_rows; r += 2)
        for (uint32_t c = 0; c < ctx->mi_cols; c += 2)
            cdef_block(ctx, g, src, nsrc, dst, ndst, np, r, c, cdef_idx_at(ctx, g, r & ~15u, c & ~15u));
}

/* ===================================================================== */
/* §7.16 superres                                                         */
/* ===================================================================== */

/* upscale_sample (superres.rs:422-445). */
int32_t oo_upscale_sample(const int32_t *row, uint32_t row_len, int32_t src_x_px,
                          uint32_t subpel, int32_t min_x, int32_t max_x, uint8_t bit_depth) {
    const int16_t *taps = &oo_upscale_filter[(subpel & 63) * 8];
    int32_t sum = 0;
    for (int k = 0; k < 8; k++) {
        int32_t sxp = clip3(min_x, max_x, src_x_px + (k - 3));
        int32_t idx = clip3(0, (int32_t)row_len - 1, sxp);
        sum += row[idx] * (int32_t)taps[k];
    }
    return clip3(0, (1 << bit_depth) - 1, round2_i32(sum, 7));
}

/* upscale_plane (superres.rs:326-410). */
static int upscale_plane(const oxav1_sr_ctx *ctx, uint8_t plane, const oxav1_plane_i32 *in,
                         oxav1_plane_i32 *out) {
    uint32_t sx = plane ? ctx->subsampling_x : 0, sy = plane ? ctx->subsampling_y : 0;
    uint32_t dw = round2_u32(ctx->frame_width, sx);
    uint32_t uw = round2_u32(ctx->upscaled_width, sx);
    uint32_t ph = round2_u32(ctx->frame_height, sy);
    if (in->rows < ph || in->cols < dw || out->rows != ph || out->cols != uw)
        return OXAV1_ERR_PLANE_SHAPE_MISMATCH;
    if (uw == 0 || ph == 0) return OXAV1_OK;
    if (uw <= dw) return OXAV1_ERR_UPSCALED_NOT_LARGER;
    int64_t stepx_num = ((int64_t)dw << 14) + (int64_t)(uw / 2);
    int32_t step_x = (int32_t)(stepx_num / (int64_t)uw);
    int32_t err = (int32_t)((int64_t)uw * (int64_t)step_x - ((int64_t)dw << 14));
    int64_t diff = (int64_t)uw - (int64_t)dw;
    int64_t num = -(diff << 13) + (int64_t)(uw / 2);
    int32_t division = (int32_t)(num / (int64_t)uw);
    int32_t isx = division + (1 << 7) - err / 2;
    isx &= (1 << 14) - 1;
    uint32_t mi_w = ctx->mi_cols >> sx;
    int32_t min_x = 0, max_x = (int32_t)mi_w * MI_SIZE - 1;
    for (uint32_t y = 0; y < ph; y++) {
        const int32_t *row = in->samples + (size_t)y * in->cols;
        int32_t *orow = out->samples + (size_t)y * out->cols;
        for (uint32_t x = 0; x < uw; x++) {
            int32_t src_x = -(1 << 14) + isx + (int32_t)x * step_x;
            int32_t px = src_x >> 14;
            int32_t sub = (src_x & ((1 << 14) - 1)) >> 8;
            orow[x] = oo_upscale_sample(row, in->cols, px, (uint32_t)sub, min_x, max_x, ctx->bit_depth);
        }
    }
    return OXAV1_OK;
}

/* upscale_frame (superres.rs:282-318) + copy_plane (superres.rs:505-518). */
int oo_upscale_frame(const oxav1_sr_ctx *ctx, const oxav1_plane_i32 *in, uint32_t nin,
                     oxav1_plane_i32 *out, uint32_t nout) {
    uint32_t np = ctx->num_planes;
    if (nin < np || nout < np) return OXAV1_ERR_PLANE_COUNT_MISMATCH;
    if (!ctx->use_superres || ctx->frame_width == ctx->upscaled_width) {
        for (uint32_t p = 0; p < np; p++) {
            if (in[p].rows != out[p].rows || in[p].cols != out[p].cols)
                return OXAV1_ERR_PLANE_SHAPE_MISMATCH;
            memcpy(out[p].samples, in[p].samples, (size_t)in[p].rows * in[p].cols * sizeof(int32_t));
        }
        return OXAV1_OK;
    }
    if (ctx->upscaled_width <= ctx->frame_width) return OXAV1_ERR_UPSCALED_NOT_LARGER;
    for (uint32_t p = 0; p < np; p++) {
        int st = upscale_plane(ctx, (uint8_t)p, &in[p], &out[p]);
        if (st != OXAV1_OK) return st;
    }
    return OXAV1_OK;
}

/* ===================================================================== */
/* §7.17 loop restoration                                                 */
/* ===================================================================== */

/* count_units_in_frame (loop_restoration.rs:910-915). */
uint32_t oo_count_units_in_frame(uint32_t unit_size, uint32_t frame_size) {
    if (unit_size == 0) return 1;
    uint32_t v = (frame_size + (unit_size >> 1)) / unit_size;
    return v > 1 ? v : 1;
}

static uint32_t lr_unit_size(const oxav1_lr_ctx *ctx, uint8_t plane) {
    uint32_t u = plane < 3 ? ctx->loop_restoration_size[plane] : 0;
    return u > 1 ? u : 1;
}

/* derive_block_geometry (loop_restoration.rs:370-423). */
oo_lr_geom oo_lr_derive_block_geometry(const oxav1_lr_ctx *ctx, uint8_t plane, uint32_t row,
                                       uint32_t col) {
    oo_lr_geom g;
    int32_t luma_y = (int32_t)row * MI_SIZE;
    int32_t stripe_num = (luma_y + 8) / 64;
    uint32_t sx = plane ? ctx->subsampling_x : 0, sy = plane ? ctx->subsampling_y : 0;
    g.stripe_start_y = (-8 + stripe_num * 64) >> sy;
    g.stripe_end_y = g.stripe_start_y + (64 >> sy) - 1;
    uint32_t us = lr_unit_size(ctx, plane);
    uint32_t unit_rows = oo_count_units_in_frame(us, round2_u32(ctx->frame_height, sy));
    uint32_t unit_cols = oo_count_units_in_frame(us, round2_u32(ctx->upscaled_width, sx));
    uint32_t ur = ((row * MI_SIZE + 8) >> sy) / us;
    g.unit_row = ur < unit_rows - 1 ? ur : unit_rows - 1;
    uint32_t uc = ((col * MI_SIZE) >> sx) / us;
    g.unit_col = uc < unit_cols - 1 ? uc : unit_cols - 1;
    g.plane_end_x = (int32_t)round2_u32(ctx->upscaled_width, sx) - 1;
    g.plane_end_y = (int32_t)round2_u32(ctx->frame_height, sy) - 1;
    g.x = (col * MI_SIZE) >> sx;
    g.y = (row * MI_SIZE) >> sy;
    int32_t wr = imax(g.plane_end_x + 1 - (int32_t)g.x, 0);
    int32_t hr = imax(g.plane_end_y + 1 - (int32_t)g.y, 0);
    g.w = (uint32_t)imin(MI_SIZE >> sx, wr);
    g.h = (uint32_t)imin(MI_SIZE >> sy, hr);
    return g;
}

/* sample_at (loop_restoration.rs:889-902). */
static int32_t lr_sample_at(const oxav1_plane_i32 *planes, uint32_t n, uint32_t p, int32_t x,
                            int32_t y) {
    if (p >= n) return 0;
    const oxav1_plane_i32 *b = &planes[p];
    if (b->rows == 0 || b->cols == 0) return 0;
    int32_t yc = clip3(0, (int32_t)b->rows - 1, y), xc = clip3(0, (int32_t)b->cols - 1, x);
    return b->samples[(size_t)yc * b->cols + xc];
}

/* get_source_sample (loop_restoration.rs:862-883). */
int32_t oo_lr_get_source_sample(const oxav1_plane_i32 *curr, uint32_t ncurr,
                                const oxav1_plane_i32 *cdef, uint32_t ncdef, uint8_t plane,
                                int32_t x, int32_t y, const oo_lr_geom *g) {
    x = clip3(0, imax(g->plane_end_x, 0), x);
    y = clip3(0, imax(g->plane_end_y, 0), y);
    if (y < g->stripe_start_y) {
        y = imax(y, g->stripe_start_y - 2);
        return lr_sample_at(curr, ncurr, plane, x, y);
    } else if (y > g->stripe_end_y) {
        y = imin(y, g->stripe_end_y + 2);
        return lr_sample_at(curr, ncurr, plane, x, y);
    }
    return lr_sample_at(cdef, ncdef, plane, x, y);
}

/* wiener_coefficients (loop_restoration.rs:839-849). */
void oo_wiener_coefficients(const int32_t coeff[3], int32_t filter[7]) {
    for (int i = 0; i < 7; i++) filter[i] = 0;
    filter[3] = 128;
    for (int i = 0; i < 3; i++) {
        filter[i] = coeff[i];
        filter[6 - i] = coeff[i];
        filter[3] -= 2 * coeff[i];
    }
}

static const oxav1_lr_unit LR_UNIT_NONE = {0, 0, {{0, 0, 0}, {0, 0, 0}}, {0, 0}};

/* lr_unit closure over the unit grid (cdf.rs:20273-20282). */
static oxav1_lr_unit lr_unit_at(const oxav1_lr_ctx *ctx, const oxav1_lr_grids *g, uint8_t plane,
                                uint32_t ur, uint32_t uc) {
    if (plane >= 3 || !g->units[plane]) return LR_UNIT_NONE;
    uint32_t sx = plane ? ctx->subsampling_x : 0, sy = plane ? ctx->subsampling_y : 0;
    uint32_t us = lr_unit_size(ctx, plane);
    uint32_t rows = oo_count_units_in_frame(us, round2_u32(ctx->frame_height, sy));
    uint32_t cols = oo_count_units_in_frame(us, round2_u32(ctx->upscaled_width, sx));
    if (ur >= rows || uc >= cols) return LR_UNIT_NONE;
    return g->units[plane][(size_t)ur * cols + uc];
}

/* box_filter (loop_restoration.rs:594-729). */
void oo_lr_box_filter(const oxav1_plane_i32 *curr, uint32_t ncurr, const oxav1_plane_i32 *cdef,
                      uint32_t ncdef, uint8_t plane, const oo_lr_geom *geom, uint8_t bit_depth,
                      int32_t r, int32_t eps, uint8_t pass, int32_t *f) {
    uint32_t h = geom->h, w = geom->w;
    memset(f, 0, sizeof(int32_t) * h * w);
    if (r == 0) return;
    int32_t n = (2 * r + 1) * (2 * r + 1);
    int64_t n2e = (int64_t)n * n * eps;
    int64_t s_div = (((int64_t)1 << 20) + n2e / 2) / (n2e > 1 ? n2e : 1);
    int32_t one_over_n = ((1 << 12) + n / 2) / n;
    uint32_t aw = w + 2, ah = h + 2;
    int32_t a_arr[6 * 6], b_arr[6 * 6];
    int32_t bd8 = imax((int32_t)bit_depth - 8, 0);
    uint32_t sh_a = (uint32_t)(2 * bd8), sh_b = (uint32_t)bd8;
    for (int32_t i = -1; i <= (int32_t)h; i++) {
        for (int32_t j = -1; j <= (int32_t)w; j++) {
            int64_t a_acc = 0, b_acc = 0;
            for (int32_t dy = -r; dy <= r; dy++)
                for (int32_t dx = -r; dx <= r; dx++) {
                    int64_t c = oo_lr_get_source_sample(curr, ncurr, cdef, ncdef, plane,
                                                        (int32_t)geom->x + j + dx,
                                                        (int32_t)geom->y + i + dy, geom);
                    a_acc += c * c;
                    b_acc += c;
                }
            int32_t a_r = (int32_t)round2_i64(a_acc, sh_a);
            int32_t d_r = (int32_t)round2_i64(b_acc, sh_b);
            int64_t p = (int64_t)a_r * n - (int64_t)d_r * d_r;
            if (p < 0) p = 0;
            int32_t z = (int32_t)round2_i64(p * s_div, 20);
            int32_t a2 = z >= 255 ? 256 : (z == 0 ? 1 : ((z << 8) + z / 2) / (z + 1));
            int64_t b2 = (int64_t)((1 << 8) - a2) * b_acc * (int64_t)one_over_n;
            size_t ai = (size_t)(i + 1), aj = (size_t)(j + 1);
            a_arr[ai * aw + aj] = a2;
            b_arr[ai * aw + aj] = (int32_t)round2_i64(b2, 12);
        }
    }
    (void)ah;
    if (plane >= ncdef) return;
    const oxav1_plane_i32 *cd = &cdef[plane];
    for (uint32_t i = 0; i < h; i++) {
        uint32_t shift = (pass == 0 && (i & 1) == 1) ? 4 : 5;
        for (uint32_t j = 0; j < w; j++) {
            int64_t a_sum = 0, b_sum = 0;
            for (int32_t dy = -1; dy <= 1; dy++)
                for (int32_t dx = -1; dx <= 1; dx++) {
                    int32_t weight;
                    if (pass == 0)
                        weight = (((int32_t)i + dy) & 1) ? (dx == 0 ? 6 : 5) : 0;
                    else
                        weight = (dx == 0 || dy == 0) ? 4 : 3;
                    size_t ai = (size_t)((int32_t)i + dy + 1), aj = (size_t)((int32_t)j + dx + 1);
                    a_sum += (int64_t)weight * a_arr[ai * aw + aj];
                    b_sum += (int64_t)weight * b_arr[ai * aw + aj];
                }
            int32_t py = (int32_t)geom->y + (int32_t)i, px = (int32_t)geom->x + (int32_t)j;
            int64_t u = (py >= 0 && py < (int32_t)cd->rows && px >= 0 && px < (int32_t)cd->cols)
                            ? cd->samples[(size_t)py * cd->cols + px]
                            : 0;
            int64_t v = a_sum * u + b_sum;
            f[i * w + j] = (int32_t)round2_i64(v, 8 + shift - 4);
        }
    }
}

/* self_guided_filter (loop_restoration.rs:485-579). */
static void lr_self_guided(const oxav1_lr_ctx *ctx, const oxav1_lr_unit *u,
                           const oxav1_plane_i32 *curr, uint32_t ncurr, const oxav1_plane_i32 *cdef,
                           uint32_t ncdef, oxav1_plane_i32 *lr, uint32_t nlr, uint8_t plane,
                           const oo_lr_geom *geom) {
    if (geom->w == 0 || geom->h == 0) return;
    uint32_t set = u->sgr_set < 15 ? u->sgr_set : 15;
    int32_t r0 = oo_sgr_params[set * 4 + 0], e0 = oo_sgr_params[set * 4 + 1];
    int32_t r1 = oo_sgr_params[set * 4 + 2], e1 = oo_sgr_params[set * 4 + 3];
    int32_t w0 = u->sgr_xqd[0], w1 = u->sgr_xqd[1];
    int32_t w2 = (1 << 7) - w0 - w1;
    int32_t flt0[16], flt1[16];
    oo_lr_box_filter(curr, ncurr, cdef, ncdef, plane, geom, ctx->bit_depth, r0, e0, 0, flt0);
    oo_lr_box_filter(curr, ncurr, cdef, ncdef, plane, geom, ctx->bit_depth, r1, e1, 1, flt1);
    uint32_t h = geom->h, w = geom->w;
    int32_t maxs = (1 << ctx->bit_depth) - 1;
    if (plane >= nlr || plane >= ncdef) return;
    oxav1_plane_i32 *out = &lr[plane];
    const oxav1_plane_i32 *cd = &cdef[plane];
    for (uint32_t i = 0; i < h; i++)
        for (uint32_t j = 0; j < w; j++) {
            int32_t py = (int32_t)geom->y + (int32_t)i, px = (int32_t)geom->x + (int32_t)j;
            int32_t uu = (py >= 0 && py < (int32_t)cd->rows && px >= 0 && px < (int32_t)cd->cols)
                             ? cd->samples[(size_t)py * cd->cols + px] << 4
                             : 0;
            int32_t v = w1 * uu;
            v += r0 != 0 ? w0 * flt0[i * w + j] : w0 * uu;
            v += r1 != 0 ? w2 * flt1[i * w + j] : w2 * uu;
            int32_t s = round2_i32(v, 11);
            int32_t o = clip3(0, maxs, s);
            if (py >= 0 && py < (int32_t)out->rows && px >= 0 && px < (int32_t)out->cols)
                out->samples[(size_t)py * out->cols + px] = o;
        }
}

/* wiener_filter (loop_restoration.rs:742-821). */
static void lr_wiener(const oxav1_lr_ctx *ctx, const oxav1_lr_unit *u, const oxav1_plane_i32 *curr,
                      uint32_t ncurr, const oxav1_plane_i32 *cdef, uint32_t ncdef,
                      oxav1_plane_i32 *lr, uint32_t nlr, uint8_t plane, const oo_lr_geom *geom) {
    if (plane >= ncurr || plane >= ncdef) return;
    int32_t vc[3] = {u->wiener[0][0], u->wiener[0][1], u->wiener[0][2]};
    int32_t hc[3] = {u->wiener[1][0], u->wiener[1][1], u->wiener[1][2]};
    int32_t vf[7], hf[7];
    oo_wiener_coefficients(vc, vf);
    oo_wiener_coefficients(hc, hf);
    uint32_t r0 = ctx->bit_depth == 12 ? 5 : 3;
    uint32_t r1 = ctx->bit_depth == 12 ? 9 : 11;
    int32_t bd = ctx->bit_depth;
    int32_t offset = 1 << (bd + 7 - (int32_t)r0 - 1);
    int32_t limit = (1 << (bd + 1 + 7 - (int32_t)r0)) - 1;
    uint32_t h = geom->h, w = geom->w;
    int32_t inter[10 * 4];
    for (uint32_t r = 0; r < h + 6; r++)
        for (uint32_t c = 0; c < w; c++) {
            int32_t s = 0;
            for (int t = 0; t < 7; t++) {
                int32_t sx = (int32_t)geom->x + (int32_t)c + t - 3;
                int32_t sy = (int32_t)geom->y + (int32_t)r - 3;
                s += hf[t] * oo_lr_get_source_sample(curr, ncurr, cdef, ncdef, plane, sx, sy, geom);
            }
            int32_t v = round2_i32(s, r0);
            inter[r * w + c] = clip3(-offset, limit - offset, v);
        }
    int32_t maxs = (1 << bd) - 1;
    if (plane >= nlr) return;
    oxav1_plane_i32 *out = &lr[plane];
    for (uint32_t r = 0; r < h; r++)
        for (uint32_t c = 0; c < w; c++) {
            int32_t s = 0;
            for (int t = 0; t < 7; t++) s += vf[t] * inter[(r + (uint32_t)t) * w + c];
            int32_t v = round2_i32(s, r1);
            int32_t o = clip3(0, maxs, v);
            int32_t dy = (int32_t)geom->y + (int32_t)r, dx = (int32_t)geom->x + (int32_t)c;
            if (dy < 0 || dy >= (int32_t)out->rows || dx < 0 || dx >= (int32_t)out->cols) continue;
            out->samples[(size_t)dy * out->cols + dx] = o;
        }
}

/* loop_restore_block (loop_restoration.rs:430-460). */
void oo_lr_restore_block(const oxav1_lr_ctx *ctx, const oxav1_lr_grids *g,
                             const oxav1_plane_i32 *curr, uint32_t ncurr,
                             const oxav1_plane_i32 *cdef, uint32_t ncdef, oxav1_plane_i32 *lr,
                             uint32_t nlr, uint8_t plane, uint32_t row, uint32_t col) {
    oo_lr_geom geom = oo_lr_derive_block_geometry(ctx, plane, row, col);
    if (geom.w == 0 || geom.h == 0) return;
    oxav1_lr_unit u = lr_unit_at(ctx, g, plane, geom.unit_row, geom.unit_col);
    if (u.type == OXAV1_RESTORE_WIENER)
        lr_wiener(ctx, &u, curr, ncurr, cdef, ncdef, lr, nlr, plane, &geom);
    else if (u.type == OXAV1_RESTORE_SGRPROJ)
        lr_self_guided(ctx, &u, curr, ncurr, cdef, ncdef, lr, nlr, plane, &geom);
}

/* loop_restoration_frame (loop_restoration.rs:268-328). */
void oo_loop_restoration_frame(const oxav1_lr_ctx *ctx, const oxav1_lr_grids *g,
                               const oxav1_plane_i32 *curr, uint32_t ncurr,
                               const oxav1_plane_i32 *cdef, uint32_t ncdef,
                               oxav1_plane_i32 *lr, uint32_t nlr) {
    uint32_t np = ctx->num_planes;
    if (np > ncurr) np = ncurr;
    if (np > ncdef) np = ncdef;
    if (np > nlr) np = nlr;
    for (uint32_t p = 0; p < np; p++) {
        uint32_t rows = cdef[p].rows < lr[p].rows ? cdef[p].rows : lr[p].rows;
        uint32_t cols = cdef[p].cols < lr[p].cols ? cdef[p].cols : lr[p].cols;
        for (uint32_t y = 0; y < rows; y++)
            for (uint32_t x = 0; x < cols; x++)
                lr[p].samples[(size_t)y * lr[p].cols + x] = cdef[p].samples[(size_t)y * cdef[p].cols + x];
    }
    if (!ctx->uses_lr) return;
    for (uint32_t y = 0; y < ctx->frame_height; y += MI_SIZE)
        for (uint32_t x = 0; x < ctx->upscaled_width; x += MI_SIZE)
            for (uint32_t p = 0; p < np; p++) {
                if (ctx->frame_restoration_type[p] == OXAV1_RESTORE_NONE) continue;
                oo_lr_restore_block(ctx, g, curr, ncurr, cdef, ncdef, lr, nlr, (uint8_t)p, y >> 2, x >> 2);
            }
}

/* ===================================================================== */
/* §7.18.3 film grain                                                     */
/* ===================================================================== */

/* RandomRegister::next (film_grain.rs:117-130). */
uint32_t oo_fg_random(uint16_t *state, uint32_t bits) {
    uint32_t r = *state;
    uint32_t bit = (r ^ (r >> 1) ^ (r >> 3) ^ (r >> 12)) & 1;
    r = (r >> 1) | (bit << 15);
    *state = (uint16_t)r;
    return (r >> (16 - bits)) & ((1u << bits) - 1);
}

#define LGW 82
#define LGH 73

/* generate_grain (film_grain.rs:230-382). */
uint32_t oo_fg_generate_grain(const oxav1_film_grain_params *fg, uint8_t bd, uint8_t sx,
                              uint8_t sy, int32_t *luma, int32_t *cb, int32_t *cr) {
    int32_t gc = 128 << (bd - 8);
    int32_t gmin = -gc, gmax = (256 << (bd - 8)) - 1 - gc;
    uint32_t init_shift = (12u - (uint32_t)bd) + fg->grain_scale_shift;
    uint16_t rr = fg->grain_seed;
    for (int y = 0; y < LGH; y++)
        for (int x = 0; x < LGW; x++) {
            int32_t gv = 0;
            if (fg->num_y_points > 0) gv = oo_gaussian_sequence[oo_fg_random(&rr, 11)];
            luma[y * LGW + x] = round2_i32(gv, init_shift);
        }
    if (fg->num_y_points > 0) {
        uint32_t sh = fg->ar_coeff_shift;
        int32_t lag = fg->ar_coeff_lag;
        for (int32_t y = 3; y < LGH; y++)
            for (int32_t x = 3; x < LGW - 3; x++) {
                int32_t s = 0;
                int pos = 0;
                for (int32_t dr = -lag; dr <= 0; dr++) {
                    int done = 0;
                    for (int32_t dc = -lag; dc <= lag; dc++) {
                        if (dr == 0 && dc == 0) {
                            done = 1;
                            break;
                        }
                        int32_t c = (int32_t)fg->ar_coeffs_y_plus_128[pos] - 128;
                        s += luma[(y + dr) * LGW + (x + dc)] * c;
                        pos++;
                    }
                    if (done) break;
                }
                int32_t v = luma[y * LGW + x] + round2_i32(s, sh);
                luma[y * LGW + x] = clip3(gmin, gmax, v);
            }
    }
    int32_t cw = sx ? 44 : 82, ch = sy ? 38 : 73;
    rr = (uint16_t)(fg->grain_seed ^ 0xb524);
    for (int y = 0; y < ch; y++)
        for (int x = 0; x < cw; x++) {
            int32_t gv = 0;
            if (fg->num_cb_points > 0 || fg->chroma_scaling_from_luma)
                gv = oo_gaussian_sequence[oo_fg_random(&rr, 11)];
            cb[y * cw + x] = round2_i32(gv, init_shift);
        }
    rr = (uint16_t)(fg->grain_seed ^ 0x49d8);
    for (int y = 0; y < ch; y++)
        for (int x = 0; x < cw; x++) {
            int32_t gv = 0;
            if (fg->num_cr_points > 0 || fg->chroma_scaling_from_luma)
                gv = oo_gaussian_sequence[oo_fg_random(&rr, 11)];
            cr[y * cw + x] = round2_i32(gv, init_shift);
        }
    if (fg->chroma_scaling_from_luma || fg->num_cb_points > 0 || fg->num_cr_points > 0) {
        uint32_t sh = fg->ar_coeff_shift;
        int32_t lag = fg->ar_coeff_lag;
        for (int32_t y = 3; y < ch; y++)
            for (int32_t x = 3; x < cw - 3; x++) {
                int32_t s0 = 0, s1 = 0;
                int pos = 0;
                for (int32_t dr = -lag; dr <= 0; dr++) {
                    int done = 0;
                    for (int32_t dc = -lag; dc <= lag; dc++) {
                        int32_t c0 = (int32_t)fg->ar_coeffs_cb_plus_128[pos] - 128;
                        int32_t c1 = (int32_t)fg->ar_coeffs_cr_plus_128[pos] - 128;
                        if (dr == 0 && dc == 0) {
                            if (fg->num_y_points > 0) {
                                int32_t lx = ((x - 3) << sx) + 3, ly = ((y - 3) << sy) + 3;
                                int32_t l = 0;
                                for (int32_t i = 0; i <= sy; i++)
                                    for (int32_t j = 0; j <= sx; j++) l += luma[(ly + i) * LGW + (lx + j)];
                                l = round2_i32(l, (uint32_t)sx + (uint32_t)sy);
                                s0 += l * c0;
                                s1 += l * c1;
                            }
                            done = 1;
                            break;
                        }
                        s0 += cb[(y + dr) * cw + (x + dc)] * c0;
                        s1 += cr[(y + dr) * cw + (x + dc)] * c1;
                        pos++;
                    }
                    if (done) break;
                }
                int32_t v0 = cb[y * cw + x] + round2_i32(s0, sh);
                int32_t v1 = cr[y * cw + x] + round2_i32(s1, sh);
                cb[y * cw + x] = clip3(gmin, gmax, v0);
                cr[y * cw + x] = clip3(gmin, gmax, v1);
            }
    }
    return (uint32_t)cw | ((uint32_t)ch << 16);
}

/* scaling_lookup_fill (film_grain.rs:480-526). */
static void fg_lut_fill(uint8_t *t, uint32_t np, const uint8_t *xs, const uint8_t *ys) {
    if (np == 0) {
        memset(t, 0, 256);
        return;
    }
    uint32_t x0 = xs[0];
    for (uint32_t v = 0; v < (x0 < 256 ? x0 : 256); v++) t[v] = ys[0];
    for (uint32_t i = 0; i + 1 < np; i++) {
        int32_t xa = xs[i], xb = xs[i + 1], ya = ys[i], yb = ys[i + 1];
        int32_t dx = xb - xa;
        if (dx <= 0) continue;
        int32_t dy = yb - ya;
        int32_t delta = dy * ((65536 + (dx >> 1)) / dx);
        for (int32_t x = 0; x < dx; x++) {
            int32_t v = ya + ((x * delta + 32768) >> 16);
            int32_t idx = xa + x;
            if (idx < 256) t[idx] = (uint8_t)clip3(0, 255, v);
        }
    }
    uint32_t xn = xs[np - 1];
    for (uint32_t v = (xn < 256 ? xn : 256); v < 256; v++) t[v] = ys[np - 1];
}

/* scaling_lookup_init (film_grain.rs:418-466). */
void oo_fg_scaling_lookup_init(const oxav1_film_grain_params *fg, uint8_t num_planes,
                               uint8_t lut[3][256]) {
    memset(lut, 0, 3 * 256);
    uint32_t np = num_planes < 3 ? num_planes : 3;
    for (uint32_t p = 0; p < np; p++) {
        if (p == 0 || fg->chroma_scaling_from_luma)
            fg_lut_fill(lut[p], fg->num_y_points, fg->point_y_value, fg->point_y_scaling);
        else if (p == 1)
            fg_lut_fill(lut[p], fg->num_cb_points, fg->point_cb_value, fg->point_cb_scaling);
        else
            fg_lut_fill(lut[p], fg->num_cr_points, fg->point_cr_value, fg->point_cr_scaling);
    }
}

/* scale_lut (film_grain.rs:543-555). */
int32_t oo_fg_scale_lut(const uint8_t lut[3][256], uint8_t plane, int32_t index, uint8_t bd) {
    uint32_t shift = (uint32_t)bd - 8;
    uint32_t x = (uint32_t)(index >> shift);
    const uint8_t *t = lut[plane];
    if (bd == 8 || x >= 255) return t[x < 255 ? x : 255];
    int32_t rem = index - ((int32_t)x << shift);
    int32_t start = t[x], end = t[x + 1];
    return start + round2_i32((end - start) * rem, shift);
}

/* add_noise_synthesis (film_grain.rs:570-829). */
static void fg_add_noise(const oxav1_film_grain_params *fg, const int32_t *gl, const int32_t *gcb,
                         const int32_t *gcr, int32_t cw, const uint8_t lut[3][256], uint8_t bd,
                         uint8_t num_planes, uint8_t sx, uint8_t sy, uint8_t mc,
                         oxav1_plane_i32 *out, uint32_t nout) {
    if (nout == 0) return;
    uint32_t w = out[0].cols, h = out[0].rows;
    if (w == 0 || h == 0) return;
    int32_t gc = 128 << (bd - 8);
    int32_t gmin = -gc, gmax = (256 << (bd - 8)) - 1 - gc;
    uint32_t np = num_planes < 3 ? num_planes : 3;
    uint32_t hh = (h + 1) / 2, wh = (w + 1) / 2;
    uint32_t nstripes = (hh + 15) / 16;
    size_t sh_[3] = {0, 0, 0}, sw_[3] = {0, 0, 0};
    for (uint32_t p = 0; p < np; p++) {
        uint32_t psx = p ? sx : 0, psy = p ? sy : 0;
        sh_[p] = 34u >> psy;
        sw_[p] = (size_t)((w + psx) >> psx) + 34;
    }
    int32_t *stripes[3] = {0, 0, 0};
    for (uint32_t p = 0; p < np; p++)
        stripes[p] = (int32_t *)calloc((size_t)nstripes * sh_[p] * sw_[p], sizeof(int32_t));
#define NS(p, s, i, x) stripes[p][((size_t)(s) * sh_[p] + (size_t)(i)) * sw_[p] + (size_t)(x)]
    uint32_t luma_num = 0;
    for (uint32_t ys = 0; ys < hh; ys += 16, luma_num++) {
        uint16_t rr = fg->grain_seed;
        uint32_t xor_hi = ((luma_num * 37u + 178u) & 255u) << 8;
        uint32_t xor_lo = (luma_num * 173u + 105u) & 255u;
        rr ^= (uint16_t)xor_hi;
        rr ^= (uint16_t)xor_lo;
        for (uint32_t xb = 0; xb < wh; xb += 16) {
            uint32_t rnd = oo_fg_random(&rr, 8);
            int32_t ox = (int32_t)(rnd >> 4), oy = (int32_t)(rnd & 15);
            for (uint32_t p = 0; p < np; p++) {
                uint32_t psx = p ? sx : 0, psy = p ? sy : 0;
                int32_t pox = psx ? 6 + ox : 9 + ox * 2;
                int32_t poy = psy ? 6 + oy : 9 + oy * 2;
                int32_t hs = 34 >> psy, ws = 34 >> psx;
                for (int32_t i = 0; i < hs; i++)
                    for (int32_t j = 0; j < ws; j++) {
                        int32_t yi = poy + i, xi = pox + j;
                        int32_t g = p == 0 ? gl[yi * LGW + xi] : (p == 1 ? gcb[yi * cw + xi] : gcr[yi * cw + xi]);
                        if (psx == 0) {
                            int32_t col = (int32_t)xb * 2 + j;
                            if (j < 2 && fg->overlap_flag && xb > 0) {
                                int32_t old = NS(p, luma_num, i, col);
                                g = j == 0 ? old * 27 + g * 17 : old * 17 + g * 27;
                                g = clip3(gmin, gmax, round2_i32(g, 5));
                            }
                            if ((size_t)col < sw_[p]) NS(p, luma_num, i, col) = g;
                        } else {
                            int32_t col = (int32_t)xb + j;
                            if (j == 0 && fg->overlap_flag && xb > 0) {
                                int32_t old = NS(p, luma_num, i, col);
                                g = old * 23 + g * 22;
                                g = clip3(gmin, gmax, round2_i32(g, 5));
                            }
                            if ((size_t)col < sw_[p]) NS(p, luma_num, i, col) = g;
                        }
                    }
            }
        }
    }
    int32_t *img[3] = {0, 0, 0};
    for (uint32_t p = 0; p < np; p++) {
        uint32_t psx = p ? sx : 0, psy = p ? sy : 0;
        uint32_t ph = (h + psy) >> psy, pw = (w + psx) >> psx;
        img[p] = (int32_t *)calloc((size_t)ph * pw, sizeof(int32_t));
        for (uint32_t yy = 0; yy < ph; yy++) {
            uint32_t si = yy >> (5 - psy);
            uint32_t i = yy - (si << (5 - psy));
            if (si >= nstripes) continue;
            for (uint32_t xx = 0; xx < pw; xx++) {
                int32_t g = NS(p, si, i, xx);
                if (psy == 0) {
                    if (i < 2 && si > 0 && fg->overlap_flag) {
                        int32_t old = NS(p, si - 1, i + 32, xx);
                        g = i == 0 ? old * 27 + g * 17 : old * 17 + g * 27;
                        g = clip3(gmin, gmax, round2_i32(g, 5));
                    }
                } else if (i < 1 && si > 0 && fg->overlap_flag) {
                    int32_t old = NS(p, si - 1, i + 16, xx);
                    g = old * 23 + g * 22;
                    g = clip3(gmin, gmax, round2_i32(g, 5));
                }
                img[p][(size_t)yy * pw + xx] = g;
            }
        }
    }
#undef NS
    int32_t min_v, max_l, max_c;
    if (fg->clip_to_restricted_range) {
        min_v = 16 << (bd - 8);
        max_l = 235 << (bd - 8);
        max_c = mc == 0 ? max_l : (240 << (bd - 8));
    } else {
        min_v = 0;
        max_l = max_c = (256 << (bd - 8)) - 1;
    }
    uint32_t ssh = fg->grain_scaling;
    if (np >= 3 && nout >= 3 &&
        (fg->num_cb_points > 0 || fg->num_cr_points > 0 || fg->chroma_scaling_from_luma)) {
        uint32_t pwc = (w + sx) >> sx, phc = (h + sy) >> sy;
        const oxav1_plane_i32 *oy_ = &out[0];
        for (uint32_t yy = 0; yy < phc; yy++)
            for (uint32_t xx = 0; xx < pwc; xx++) {
                uint32_t lx = xx << sx, ly = yy << sy;
                uint32_t lnx = lx + 1 < w - 1 ? lx + 1 : w - 1;
                int32_t avg = sx ? round2_i32(oy_->samples[(size_t)ly * oy_->cols + lx] +
                                                  oy_->samples[(size_t)ly * oy_->cols + lnx],
                                              1)
                                 : oy_->samples[(size_t)ly * oy_->cols + lx];
                if (fg->num_cb_points > 0 || fg->chroma_scaling_from_luma) {
                    int32_t *o = &out[1].samples[(size_t)yy * out[1].cols + xx];
                    int32_t orig = *o, merged;
                    if (fg->chroma_scaling_from_luma) {
                        merged = avg;
                    } else {
                        int32_t comb = avg * ((int32_t)fg->cb_luma_mult - 128) + orig * ((int32_t)fg->cb_mult - 128);
                        merged = clip3(0, (1 << bd) - 1, (comb >> 6) + (((int32_t)fg->cb_offset - 256) * (1 << (bd - 8))));
                    }
                    int32_t nz = img[1][(size_t)yy * pwc + xx];
                    nz = round2_i32(oo_fg_scale_lut(lut, 1, merged, bd) * nz, ssh);
                    *o = clip3(min_v, max_c, orig + nz);
                }
                if (fg->num_cr_points > 0 || fg->chroma_scaling_from_luma) {
                    int32_t *o = &out[2].samples[(size_t)yy * out[2].cols + xx];
                    int32_t orig = *o, merged;
                    if (fg->chroma_scaling_from_luma) {
                        merged = avg;
                    } else {
                        int32_t comb = avg * ((int32_t)fg->cr_luma_mult - 128) + orig * ((int32_t)fg->cr_mult - 128);
                        merged = clip3(0, (1 << bd) - 1, (comb >> 6) + (((int32_t)fg->cr_offset - 256) * (1 << (bd - 8))));
                    }
                    int32_t nz = img[2][(size_t)yy * pwc + xx];
                    nz = round2_i32(oo_fg_scale_lut(lut, 2, merged, bd) * nz, ssh);
                    *o = clip3(min_v, max_c, orig + nz);
                }
            }
    }
    if (fg->num_y_points > 0) {
        for (uint32_t yy = 0; yy < h; yy++)
            for (uint32_t xx = 0; xx < w; xx++) {
                int32_t *o = &out[0].samples[(size_t)yy * out[0].cols + xx];
                int32_t orig = *o;
                int32_t nz = img[0][(size_t)yy * w + xx];
                nz = round2_i32(oo_fg_scale_lut(lut, 0, orig, bd) * nz, ssh);
                *o = clip3(min_v, max_l, orig + nz);
            }
    }
    for (uint32_t p = 0; p < np; p++) {
        free(stripes[p]);
        free(img[p]);
    }
}

/* film_grain_synthesis (film_grain.rs:848-873). */
void oo_film_grain_synthesis(const oxav1_film_grain_params *fg, uint8_t bd, uint8_t num_planes,
                             uint8_t sx, uint8_t sy, uint8_t mc, oxav1_plane_i32 *planes,
                             uint32_t n) {
    if (!fg->apply_grain) return;
    int32_t luma[LGH * LGW], cb[LGH * LGW], cr[LGH * LGW];
    uint32_t dims = oo_fg_generate_grain(fg, bd, sx, sy, luma, cb, cr);
    uint8_t lut[3][256];
    oo_fg_scaling_lookup_init(fg, num_planes, lut);
    fg_add_noise(fg, luma, cb, cr, (int32_t)(dims & 0xffff), (const uint8_t(*)[256])lut, bd,
                 num_planes, sx, sy, mc, planes, n);
}

/* ===================================================================== */
/* Chain (frame_driver.rs:790-1116)                                       */
/* ===================================================================== */

static double now_ns(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec * 1e9 + (double)ts.tv_nsec;
}

static int32_t *dup_plane(const oxav1_plane_i32 *p) {
    size_t n = (size_t)p->rows * p->cols;
    int32_t *d = (int32_t *)malloc(n * sizeof(int32_t) + 4);
    memcpy(d, p->samples, n * sizeof(int32_t));
    return d;
}

int oo_chain(const oxav1_frame_desc *d, const oo_chain_grids *g, oxav1_plane_i32 *rec,
             oxav1_plane_i32 *ref_out, oxav1_plane_i32 *out, double *stage_ns) {
    uint32_t np = d->num_planes;
    uint32_t sx = d->subsampling_x, sy = d->subsampling_y;
    double t0 = now_ns(), t;
    double st[5] = {0, 0, 0, 0, 0};
    /* §7.14 deblock (frame_driver.rs:832-858). */
    if (d->apply_lf) oo_loop_filter_frame(&d->lf, &g->lf, rec, np);
    t = now_ns();
    st[0] = t - t0;
    t0 = t;
    /* deblocked copy for LR (864-868). */
    int uses_lr = d->lr.uses_lr;
    oxav1_plane_i32 deb[3], cur[3];
    for (uint32_t p = 0; p < np; p++) {
        cur[p] = rec[p];
        deb[p] = rec[p];
        if (uses_lr) deb[p].samples = dup_plane(&rec[p]);
    }
    /* §7.15 CDEF (869-900): src = clone, dst = live. */
    if (d->apply_cdef) {
        oxav1_plane_i32 src[3];
        for (uint32_t p = 0; p < np; p++) {
            src[p] = rec[p];
            src[p].samples = dup_plane(&rec[p]);
        }
        oxav1_cdef_ctx cc;
        cc.mi_rows = d->mi_rows;
        cc.mi_cols = d->mi_cols;
        cc.num_planes = d->num_planes;
        cc.bit_depth = d->bit_depth;
        cc.subsampling_x = d->subsampling_x;
        cc.subsampling_y = d->subsampling_y;
        cc.params = d->cdef;
        oo_cdef_frame(&cc, &g->cdef, src, np, cur, np);
        for (uint32_t p = 0; p < np; p++) free(src[p].samples);
    }
    t = now_ns();
    st[1] = t - t0;
    t0 = t;
    /* §7.16 superres (901-966), applied to the CDEF set and the deblocked set. */
    int sr = d->use_superres && d->upscaled_width > d->frame_width;
    oxav1_plane_i32 upc[3], upd[3];
    int own_upc = 0, own_upd = 0;
    if (sr) {
        oxav1_sr_ctx s;
        memset(&s, 0, sizeof s);
        s.use_superres = 1;
        s.num_planes = d->num_planes;
        s.bit_depth = d->bit_depth;
        s.subsampling_x = d->subsampling_x;
        s.subsampling_y = d->subsampling_y;
        s.frame_width = d->frame_width;
        s.upscaled_width = d->upscaled_width;
        s.frame_height = d->frame_height;
        s.mi_cols = d->mi_cols;
        for (uint32_t p = 0; p < np; p++) {
            uint32_t ow = p ? (d->upscaled_width + sx) >> sx : d->upscaled_width;
            uint32_t oh = p ? (d->frame_height + sy) >> sy : d->frame_height;
            upc[p].rows = upd[p].rows = oh;
            upc[p].cols = upd[p].cols = ow;
            upc[p].samples = (int32_t *)calloc((size_t)ow * oh, sizeof(int32_t));
            upd[p].samples = uses_lr ? (int32_t *)calloc((size_t)ow * oh, sizeof(int32_t)) : NULL;
        }
        own_upc = 1;
        int st_ = oo_upscale_frame(&s, cur, np, upc, np);
        if (st_ == OXAV1_OK && uses_lr) {
            own_upd = 1;
            st_ = oo_upscale_frame(&s, deb, np, upd, np);
        }
        if (st_ != OXAV1_OK) {
            for (uint32_t p = 0; p < np; p++) {
                free(upc[p].samples);
                free(upd[p].samples);
                if (uses_lr) free(deb[p].samples);
            }
            return st_;
        }
    } else {
        for (uint32_t p = 0; p < np; p++) {
            upc[p] = cur[p];
            upd[p] = deb[p];
        }
    }
    t = now_ns();
    st[2] = t - t0;
    t0 = t;
    /* §7.17 LR (968-1011): curr = deblocked, cdef = clone of CDEF output, out = live. */
    oxav1_plane_i32 fin[3];
    int own_fin = 0;
    if (uses_lr) {
        oxav1_plane_i32 cdc[3];
        for (uint32_t p = 0; p < np; p++) {
            cdc[p] = upc[p];
            cdc[p].samples = dup_plane(&upc[p]);
        }
        oo_loop_restoration_frame(&d->lr, &g->lr, upd, np, cdc, np, upc, np);
        for (uint32_t p = 0; p < np; p++) free(cdc[p].samples);
    }
    for (uint32_t p = 0; p < np; p++) fin[p] = upc[p];
    (void)own_fin;
    t = now_ns();
    st[3] = t - t0;
    t0 = t;
    /* crop (1013-1042) into ref_out, then FG on out (1054-1096). */
    for (uint32_t p = 0; p < np; p++) {
        uint32_t ow = p ? (d->upscaled_width + sx) >> sx : d->upscaled_width;
        uint32_t oh = p ? (d->frame_height + sy) >> sy : d->frame_height;
        for (uint32_t y = 0; y < oh; y++) {
            memcpy(ref_out[p].samples + (size_t)y * ref_out[p].cols,
                   fin[p].samples + (size_t)y * fin[p].cols, ow * sizeof(int32_t));
            memcpy(out[p].samples + (size_t)y * out[p].cols,
                   fin[p].samples + (size_t)y * fin[p].cols, ow * sizeof(int32_t));
        }
    }
    if (d->fg.apply_grain)
        oo_film_grain_synthesis(&d->fg, d->bit_depth, d->num_planes, d->subsampling_x,
                                d->subsampling_y, d->matrix_coefficients, out, np);
    t = now_ns();
    st[4] = t - t0;
    for (uint32_t p = 0; p < np; p++) {
        if (own_upc) free(upc[p].samples);
        if (own_upd) free(upd[p].samples);
        if (uses_lr) free(deb[p].samples);
    }
    if (stage_ns)
        for (int i = 0; i < 5; i++) stage_ns[i] = st[i];
    return OXAV1_OK;
}

/* ===================================================================== */
/* Table checksums                                                        */
/* ===================================================================== */
#define CK(arr)                                                     \
    do {                                                            \
        int64_t s = 0;                                              \
        for (size_t i = 0; i < sizeof(arr) / sizeof(arr[0]); i++)   \
            s += (int64_t)(i + 1) * (int64_t)arr[i];                \
        return s;                                                   \
    } while (0)
int64_t oo_table_checksum(const char *name) {
    if (!strcmp(name, "GAUSSIAN_SEQUENCE")) CK(oo_gaussian_sequence);
    if (!strcmp(name, "UPSCALE_FILTER")) CK(oo_upscale_filter);
    if (!strcmp(name, "SGR_PARAMS")) CK(oo_sgr_params);
    if (!strcmp(name, "CDEF_DIRECTIONS")) CK(oo_cdef_directions);
    if (!strcmp(name, "DIV_TABLE")) CK(oo_div_table);
    if (!strcmp(name, "CDEF_UV_DIR")) CK(oo_cdef_uv_dir);
    if (!strcmp(name, "NUM_4X4_BLOCKS_WIDE")) CK(oo_num4x4_w);
    if (!strcmp(name, "NUM_4X4_BLOCKS_HIGH")) CK(oo_num4x4_h);
    if (!strcmp(name, "TX_WIDTH")) CK(oo_tx_width);
    if (!strcmp(name, "TX_HEIGHT")) CK(oo_tx_height);
    if (!strcmp(name, "SUBSAMPLED_SIZE")) CK(oo_subsampled_size);
    if (!strcmp(name, "MAX_TX_SIZE_RECT")) CK(oo_max_tx_size_rect);
    return INT64_MIN;
}
