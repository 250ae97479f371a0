/*
 * oo_lr_elect.c — CPU oracle of the encoder's loop-restoration election
 * (encoder/lr_elect.rs) and of the rate model it prices units with: the
 * counting symbol writer (encoder/symbol_writer.rs:168-530) driving
 * write_lr_unit (encoder/loop_restoration_write.rs:95-203).  TEST
 * INFRASTRUCTURE ONLY (see oxav1_oracle.h).
 *
 * A literal restatement: every candidate is evaluated by copying the
 * unit's rects from the CDEF planes and running the oracle's
 * loop_restore_block over the unit's blocks with only that unit restored
 * (eval_unit, lr_elect.rs:216-281); the Wiener fit reads the CDEF plane
 * through the reference's clamped accessor, pixel by pixel in rect order,
 * in double precision with the reference's operation order (this file is
 * built with -ffp-contract=off); the self-guided weights come from the
 * (128, 0) / (0, 128) probe runs.  Only the SWITCHABLE arm of write_lr_unit
 * is restated: the election prices every unit under it.
 *
 * Parity pins (tests/test_oracle_lr_elect.py): the per-unit candidate SSDs
 * against numpy sums over the oracle's loop_restoration_frame output with
 * every unit set to that candidate; the self-guided weights against a numpy
 * least-squares solve of the probe deltas; each Wiener fit step against a
 * numpy lstsq over the explicitly built basis; the unit rects against the
 * plane partition; log2_frac256 and the symbol costs against
 * 256*log2(range/width); the subexp codes against the §5.9.28 decoder;
 * every elected unit as the first argmin of D*256 + lambda*R; the plan's d
 * against apply_lr_plan; and the reference's tripwire (tests/lr_ab.rs:
 * detail content under coarse quantisation must elect restoration).
 */
#include "oxav1_oracle.h"
#include "oo_tables.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* loop_restoration.rs:102-159 */
static const int32_t WIENER_TAPS_MID[3] = {3, -7, 15};
static const int32_t WIENER_TAPS_MIN[3] = {-5, -23, -17};
static const int32_t WIENER_TAPS_MAX[3] = {10, 8, 46};
static const uint32_t WIENER_TAPS_K[3] = {1, 2, 3};
static const int32_t SGRPROJ_XQD_MID[2] = {-32, 31};
static const int32_t SGRPROJ_XQD_MIN[2] = {-96, -32};
static const int32_t SGRPROJ_XQD_MAX[2] = {31, 95};
#define SGRPROJ_PARAMS_BITS 4
#define SGRPROJ_PRJ_SUBEXP_K 4
#define SGRPROJ_PRJ_BITS 7
#define EC_PROB_SHIFT 6
#define EC_MIN_PROB 4

/* ===================================================================== */
/* counting symbol writer (symbol_writer.rs)                              */
/* ===================================================================== */

typedef struct {
    uint32_t range;
    uint64_t renorm_bits;
    uint32_t initial_log2_frac256;
    int err;
    uint8_t *rec; /* test hook: the bools written, when set */
    uint32_t nrec, cap;
} oo_cw;

static uint32_t cw_floor_log2(uint32_t x) { return x == 0 ? 0 : 31u - (uint32_t)__builtin_clz(x); }

/* log2_frac256 (symbol_writer.rs:99-111): 8 fractional bits of log2(x/2^15)
 * by repeated squaring of the Q32 mantissa. */
uint32_t oo_log2_frac256(uint32_t x) {
    uint64_t v = (uint64_t)x << 17;
    uint32_t frac = 0;
    for (int i = 0; i < 8; i++) {
        v = (uint64_t)(((unsigned __int128)v * v) >> 32);
        frac <<= 1;
        if (v >= ((uint64_t)2 << 32)) {
            frac |= 1;
            v >>= 1;
        }
    }
    return frac;
}

/* new_counting (symbol_writer.rs:215-229) */
static void cw_init(oo_cw *w, uint32_t range) {
    w->range = range;
    w->renorm_bits = 0;
    w->initial_log2_frac256 = oo_log2_frac256(range);
    w->err = 0;
    w->rec = NULL;
    w->nrec = w->cap = 0;
}

/* cost_bits256 (symbol_writer.rs:259-265) */
static uint64_t cw_cost(const oo_cw *w) {
    int64_t cur = oo_log2_frac256(w->range);
    int64_t total = (int64_t)w->renorm_bits * 256 + (int64_t)w->initial_log2_frac256 - cur;
    return total > 0 ? (uint64_t)total : 0;
}

/* cur_for_symbol (symbol_writer.rs:534-540) */
static uint32_t cw_cur(const oo_cw *w, const uint16_t *cdf, uint32_t s, uint32_t n) {
    uint32_t f = (1u << 15) - (uint32_t)cdf[s];
    uint32_t cur = ((w->range >> 8) * (f >> EC_PROB_SHIFT)) >> (7 - EC_PROB_SHIFT);
    cur += EC_MIN_PROB * (n - s - 1);
    return cur;
}

/* write_symbol (symbol_writer.rs:273-346) in counting mode; the §8.3
 * adaptation is not restated (each price uses a fresh CDF copy and writes
 * one adaptive symbol).  Returns nonzero once the range invariant broke. */
static int cw_symbol(oo_cw *w, uint32_t symbol, const uint16_t *cdf, uint32_t n) {
    if (w->err) return w->err;
    uint32_t cur = cw_cur(w, cdf, symbol, n);
    uint32_t prev = symbol == 0 ? w->range : cw_cur(w, cdf, symbol - 1, n);
    w->range = prev - cur;
    uint32_t bits = 15 - cw_floor_log2(w->range);
    w->range <<= bits;
    w->renorm_bits += bits;
    if (w->range < (1u << 15) || w->range > (1u << 16) - 1) w->err = 1;
    return w->err;
}

/* write_bool (symbol_writer.rs:361-364) */
static int cw_bool(oo_cw *w, uint32_t bit) {
    const uint16_t cdf[3] = {1 << 14, 1 << 15, 0};
    if (w->rec && w->nrec < w->cap) w->rec[w->nrec++] = (uint8_t)bit;
    return cw_symbol(w, bit, cdf, 2);
}

/* write_literal (symbol_writer.rs:373-385) */
static int cw_literal(oo_cw *w, uint32_t n, uint32_t value) {
    for (int i = (int)n - 1; i >= 0; i--)
        if (cw_bool(w, (value >> i) & 1u)) return w->err;
    return 0;
}

/* write_ns (symbol_writer.rs:410-422) */
static int cw_ns(oo_cw *w, uint32_t n, uint32_t value) {
    uint32_t wd = cw_floor_log2(n) + 1;
    uint32_t m = (1u << wd) - n;
    if (value < m) return cw_literal(w, wd - 1, value);
    return cw_literal(w, wd, value + m);
}

/* write_subexp_bool (symbol_writer.rs:456-483) */
static int cw_subexp(oo_cw *w, uint32_t num_syms, uint32_t k, uint32_t value) {
    uint32_t i = 0, mk = 0;
    for (;;) {
        uint32_t b2 = i ? k + i - 1 : k;
        uint32_t a = 1u << b2;
        if (num_syms <= mk + 3 * a) return cw_ns(w, num_syms - mk, value - mk);
        if (value >= mk + a) {
            if (cw_literal(w, 1, 1)) return w->err;
            i++;
            mk += a;
        } else {
            if (cw_literal(w, 1, 0)) return w->err;
            return cw_literal(w, b2, value - mk);
        }
    }
}

/* recenter (symbol_writer.rs:127-135) */
int64_t oo_recenter(int64_t r, int64_t v) {
    if (v > 2 * r) return v;
    if (v >= r) return 2 * (v - r);
    return 2 * (r - v) - 1;
}

/* write_unsigned_subexp_with_ref_bool (symbol_writer.rs:497-512) */
static int cw_unsigned_ref(oo_cw *w, int64_t mx, uint32_t k, int64_t r, int64_t value) {
    int64_t v = (r << 1) <= mx ? oo_recenter(r, value) : oo_recenter(mx - 1 - r, mx - 1 - value);
    return cw_subexp(w, (uint32_t)mx, k, (uint32_t)v);
}

/* write_signed_subexp_with_ref_bool (symbol_writer.rs:520-529) */
static int cw_signed_ref(oo_cw *w, int64_t low, int64_t high, uint32_t k, int64_t r, int64_t value) {
    return cw_unsigned_ref(w, high - low, k, r - low, value - low);
}

uint64_t oo_cw_cost(uint32_t range, int kind, int64_t value, int64_t n, uint32_t k, int64_t low, int64_t r,
                    const uint16_t *cdf) {
    oo_cw w;
    cw_init(&w, range);
    switch (kind) {
        case 0: cw_symbol(&w, (uint32_t)value, cdf, (uint32_t)n); break;
        case 1: cw_literal(&w, (uint32_t)n, (uint32_t)value); break;
        case 2: cw_subexp(&w, (uint32_t)n, k, (uint32_t)value); break;
        default: cw_signed_ref(&w, low, n, k, r, value); break;
    }
    return cw_cost(&w);
}

uint64_t oo_cw_symbol_run(uint32_t range, uint32_t symbol, const uint16_t *cdf, uint32_t n, uint32_t reps) {
    oo_cw w;
    cw_init(&w, range);
    for (uint32_t i = 0; i < reps; i++) cw_symbol(&w, symbol, cdf, n);
    return cw_cost(&w);
}

uint32_t oo_cw_signed_ref_bits(int64_t low, int64_t high, uint32_t k, int64_t r, int64_t v, uint8_t *bits,
                               uint32_t cap, uint64_t *cost) {
    oo_cw w;
    cw_init(&w, 0x8000);
    w.rec = bits;
    w.cap = cap;
    cw_signed_ref(&w, low, high, k, r, v);
    if (cost) *cost = cw_cost(&w);
    return w.nrec;
}

/* ===================================================================== */
/* write_lr_unit, SWITCHABLE arm (loop_restoration_write.rs:95-203)      */
/* ===================================================================== */

typedef struct {
    int32_t ref_wiener[3][2][3];
    int32_t ref_xqd[3][2];
} oo_lr_wstate;

/* LrWriteState::new (loop_restoration_write.rs:33-40) */
static void wstate_init(oo_lr_wstate *st) {
    for (int p = 0; p < 3; p++) {
        for (int pass = 0; pass < 2; pass++)
            for (int j = 0; j < 3; j++) st->ref_wiener[p][pass][j] = WIENER_TAPS_MID[j];
        st->ref_xqd[p][0] = SGRPROJ_XQD_MID[0];
        st->ref_xqd[p][1] = SGRPROJ_XQD_MID[1];
    }
}

static int write_lr_unit(oo_cw *w, oo_lr_wstate *st, int plane, const oxav1_lr_unit *u, const uint16_t rt_cdf[4]) {
    uint32_t rt = u->type;
    if (plane >= 3 || rt > OXAV1_RESTORE_SGRPROJ) return -1;
    if (cw_symbol(w, rt, rt_cdf, 3)) return -1;
    if (rt == OXAV1_RESTORE_WIENER) {
        for (int pass = 0; pass < 2; pass++) {
            int fc = plane != 0 ? 1 : 0;
            if (plane != 0 && u->wiener[pass][0] != 0) return -1;
            for (int j = fc; j < 3; j++) {
                int64_t mn = WIENER_TAPS_MIN[j], mx = WIENER_TAPS_MAX[j];
                int64_t r = st->ref_wiener[plane][pass][j], v = u->wiener[pass][j];
                if (v < mn || v > mx) return -1;
                if (cw_signed_ref(w, mn, mx + 1, WIENER_TAPS_K[j], r, v)) return -1;
                st->ref_wiener[plane][pass][j] = (int32_t)v;
            }
        }
    } else if (rt == OXAV1_RESTORE_SGRPROJ) {
        if (u->sgr_set >= 16) return -1;
        if (cw_literal(w, SGRPROJ_PARAMS_BITS, u->sgr_set)) return -1;
        for (int i = 0; i < 2; i++) {
            int32_t radius = oo_sgr_params[u->sgr_set * 4 + i * 2];
            int64_t mn = SGRPROJ_XQD_MIN[i], mx = SGRPROJ_XQD_MAX[i], v = u->sgr_xqd[i];
            if (radius != 0) {
                if (v < mn || v > mx) return -1;
                if (cw_signed_ref(w, mn, mx + 1, SGRPROJ_PRJ_SUBEXP_K, st->ref_xqd[plane][i], v)) return -1;
            } else {
                int32_t derived = 0;
                if (i == 1) {
                    derived = (1 << SGRPROJ_PRJ_BITS) - st->ref_xqd[plane][0];
                    derived = derived < (int32_t)mn ? (int32_t)mn : (derived > (int32_t)mx ? (int32_t)mx : derived);
                }
                if (u->sgr_xqd[i] != derived) return -1;
            }
            st->ref_xqd[plane][i] = (int32_t)v;
        }
    }
    return 0;
}

/* price_unit (lr_elect.rs:537-545): a counting writer from 0x8000 and
 * copies of the CDFs and the reference state; the result of the write is
 * ignored, as the reference ignores it. */
static uint64_t price_unit(const oo_lr_wstate *st, int plane, const oxav1_lr_unit *u, const uint16_t rt_cdf[4]) {
    oo_cw w;
    cw_init(&w, 0x8000);
    oo_lr_wstate s2 = *st;
    (void)write_lr_unit(&w, &s2, plane, u, rt_cdf);
    return cw_cost(&w);
}

uint64_t oo_lr_price_unit(const int32_t ref_wiener[2][3], const int32_t ref_xqd[2], int plane,
                          const oxav1_lr_unit *u, const uint16_t rt_cdf[4]) {
    oo_lr_wstate st;
    wstate_init(&st);
    if (plane < 0 || plane > 2) return 0;
    memcpy(st.ref_wiener[plane], ref_wiener, sizeof st.ref_wiener[plane]);
    memcpy(st.ref_xqd[plane], ref_xqd, sizeof st.ref_xqd[plane]);
    return price_unit(&st, plane, u, rt_cdf);
}

/* ===================================================================== */
/* election (lr_elect.rs)                                                 */
/* ===================================================================== */

typedef struct {
    uint32_t *blocks; /* (row, col) pairs */
    uint32_t *rects;  /* (x, y, w, h) */
    uint32_t n, cap;
} oo_unit_blocks;

typedef struct {
    const oxav1_lr_elect_input *in;
    uint32_t np;
    uint32_t pw[3], ph[3];
    oxav1_lr_ctx ctx;
    oxav1_plane_i32 curr[3], cdef[3], lr[3];
    oxav1_lr_unit *grid[3]; /* eval grid: every unit NONE but the one evaluated */
    uint32_t urows[3], ucols[3];
    oo_unit_blocks *units[3];
} oo_ec;

static void ub_push(oo_unit_blocks *u, uint32_t row, uint32_t col, const oo_lr_geom *g) {
    if (u->n == u->cap) {
        u->cap = u->cap ? 2 * u->cap : 64;
        u->blocks = (uint32_t *)realloc(u->blocks, sizeof(uint32_t) * 2 * u->cap);
        u->rects = (uint32_t *)realloc(u->rects, sizeof(uint32_t) * 4 * u->cap);
    }
    u->blocks[2 * u->n] = row;
    u->blocks[2 * u->n + 1] = col;
    u->rects[4 * u->n] = g->x;
    u->rects[4 * u->n + 1] = g->y;
    u->rects[4 * u->n + 2] = g->w;
    u->rects[4 * u->n + 3] = g->h;
    u->n++;
}

/* EvalCtx + the per-plane unit maps (lr_elect.rs:425-505): the eval-side
 * header is SWITCHABLE with 64-sample units on every plane. */
static void ec_init(oo_ec *E, const oxav1_lr_elect_input *in, const uint16_t *const cdef_src[3]) {
    memset(E, 0, sizeof *E);
    E->in = in;
    E->np = in->num_planes < 3 ? in->num_planes : 3;
    oxav1_lr_ctx *c = &E->ctx;
    c->mi_rows = in->mi_rows;
    c->mi_cols = in->mi_cols;
    c->num_planes = (uint8_t)E->np;
    c->bit_depth = in->bit_depth;
    c->subsampling_x = in->subsampling_x;
    c->subsampling_y = in->subsampling_y;
    c->frame_height = in->height;
    c->upscaled_width = in->width;
    for (int p = 0; p < 3; p++) {
        c->frame_restoration_type[p] = OXAV1_RESTORE_SWITCHABLE;
        c->loop_restoration_size[p] = 64;
    }
    c->uses_lr = 1;
    for (uint32_t p = 0; p < E->np; p++) {
        E->pw[p] = p ? in->chroma_w : in->width;
        E->ph[p] = p ? in->chroma_h : in->height;
        size_t n = (size_t)E->pw[p] * E->ph[p];
        oxav1_plane_i32 *bufs[3] = {&E->curr[p], &E->cdef[p], &E->lr[p]};
        const uint16_t *srcs[3] = {in->curr[p], cdef_src[p], cdef_src[p]};
        for (int b = 0; b < 3; b++) {
            bufs[b]->rows = E->ph[p];
            bufs[b]->cols = E->pw[p];
            bufs[b]->samples = (int32_t *)malloc(sizeof(int32_t) * (n ? n : 1));
            for (size_t i = 0; i < n; i++) bufs[b]->samples[i] = srcs[b][i];
        }
        uint32_t sx = p ? in->subsampling_x : 0, sy = p ? in->subsampling_y : 0;
        E->urows[p] = oo_count_units_in_frame(64, (in->height + sy) >> sy);
        E->ucols[p] = oo_count_units_in_frame(64, (in->width + sx) >> sx);
        uint32_t nu = E->urows[p] * E->ucols[p];
        E->grid[p] = (oxav1_lr_unit *)calloc(nu, sizeof(oxav1_lr_unit));
        E->units[p] = (oo_unit_blocks *)calloc(nu, sizeof(oo_unit_blocks));
        for (uint32_t y = 0; y < in->height; y += 4)
            for (uint32_t x = 0; x < in->width; x += 4) {
                oo_lr_geom g = oo_lr_derive_block_geometry(c, (uint8_t)p, y >> 2, x >> 2);
                if (g.w > 0 && g.h > 0)
                    ub_push(&E->units[p][g.unit_row * E->ucols[p] + g.unit_col], y >> 2, x >> 2, &g);
            }
    }
}

static void ec_free(oo_ec *E) {
    for (uint32_t p = 0; p < E->np; p++) {
        free(E->curr[p].samples);
        free(E->cdef[p].samples);
        free(E->lr[p].samples);
        uint32_t nu = E->urows[p] * E->ucols[p];
        for (uint32_t u = 0; u < nu; u++) {
            free(E->units[p][u].blocks);
            free(E->units[p][u].rects);
        }
        free(E->units[p]);
        free(E->grid[p]);
    }
}

/* rect_ssd (lr_elect.rs:167-180) */
static uint64_t rect_ssd(const int32_t *a, const uint16_t *b, size_t stride, const oo_unit_blocks *ub) {
    uint64_t ssd = 0;
    for (uint32_t r = 0; r < ub->n; r++) {
        const uint32_t *q = ub->rects + 4 * r;
        for (uint32_t row = q[1]; row < q[1] + q[3]; row++)
            for (uint32_t col = q[0]; col < q[0] + q[2]; col++) {
                int64_t d = (int64_t)a[row * stride + col] - (int64_t)b[row * stride + col];
                ssd += (uint64_t)(d * d);
            }
    }
    return ssd;
}

/* eval_unit (lr_elect.rs:216-281) */
static uint64_t eval_unit(oo_ec *E, uint32_t plane, uint32_t ur, uint32_t uc, const oxav1_lr_unit *unit,
                          const oo_unit_blocks *ub) {
    size_t stride = E->pw[plane];
    for (uint32_t r = 0; r < ub->n; r++) {
        const uint32_t *q = ub->rects + 4 * r;
        for (uint32_t row = q[1]; row < q[1] + q[3]; row++)
            for (uint32_t col = q[0]; col < q[0] + q[2]; col++)
                E->lr[plane].samples[row * stride + col] = E->cdef[plane].samples[row * stride + col];
    }
    if (unit->type != OXAV1_RESTORE_NONE) {
        oxav1_lr_grids g;
        memset(&g, 0, sizeof g);
        uint32_t k = ur * E->ucols[plane] + uc;
        E->grid[plane][k] = *unit;
        g.units[plane] = E->grid[plane];
        for (uint32_t b = 0; b < ub->n; b++)
            oo_lr_restore_block(&E->ctx, &g, E->curr, E->np, E->cdef, E->np, E->lr, E->np, (uint8_t)plane,
                                ub->blocks[2 * b], ub->blocks[2 * b + 1]);
        memset(&E->grid[plane][k], 0, sizeof(oxav1_lr_unit));
    }
    return rect_ssd(E->lr[plane].samples, E->in->input[plane], stride, ub);
}

/* fit_wiener (lr_elect.rs:288-415) */
static void fit_wiener(const int32_t *cdef, const uint16_t *src, uint32_t plane_w, uint32_t plane_h,
                       const uint32_t *rects, uint32_t nrects, int first_coeff, int32_t taps[2][3],
                       double trace[4][3]) {
#define AT(xx, yy)                                                                                         \
    ((double)cdef[(size_t)((yy) < 0 ? 0 : ((yy) > (int64_t)plane_h - 1 ? (int64_t)plane_h - 1 : (yy))) * \
                      plane_w +                                                                            \
                  (size_t)((xx) < 0 ? 0 : ((xx) > (int64_t)plane_w - 1 ? (int64_t)plane_w - 1 : (xx)))])
    double vt[3] = {WIENER_TAPS_MID[0], WIENER_TAPS_MID[1], WIENER_TAPS_MID[2]};
    double ht[3];
    memcpy(ht, vt, sizeof ht);
    if (first_coeff == 1) {
        vt[0] = 0.0;
        ht[0] = 0.0;
    }
    int step = 0;
    for (int round = 0; round < 2; round++) {
        for (int dir = 0; dir < 2; dir++) {
            const double *t = dir == 0 ? vt : ht;
            double c = 128.0 - 2.0 * (t[0] + t[1] + t[2]);
            double fixed[7] = {t[0], t[1], t[2], c, t[2], t[1], t[0]};
            int nvar = 3 - first_coeff;
            double ata[3][3], atb[3];
            memset(ata, 0, sizeof ata);
            memset(atb, 0, sizeof atb);
            for (uint32_t q = 0; q < nrects; q++) {
                uint32_t rx = rects[4 * q], ry = rects[4 * q + 1], rw = rects[4 * q + 2], rh = rects[4 * q + 3];
                for (uint32_t y = ry; y < ry + rh; y++)
                    for (uint32_t x = rx; x < rx + rw; x++) {
                        double m[7];
                        for (int j = 0; j < 7; j++) {
                            int64_t off = j - 3;
                            double acc = 0.0;
                            for (int k = 0; k < 7; k++) {
                                int64_t foff = k - 3;
                                int64_t sx, sy;
                                if (dir == 0) {
                                    sx = (int64_t)x + off;
                                    sy = (int64_t)y + foff;
                                } else {
                                    sx = (int64_t)x + foff;
                                    sy = (int64_t)y + off;
                                }
                                acc += fixed[k] * AT(sx, sy);
                            }
                            m[j] = acc / 128.0;
                        }
                        double target = (double)src[(size_t)y * plane_w + x] - m[3];
                        double basis[3];
                        for (int i = 0; i < 3; i++) basis[i] = (m[i] + m[6 - i] - 2.0 * m[3]) / 128.0;
                        for (int i = first_coeff; i < 3; i++) {
                            for (int j = first_coeff; j < 3; j++) ata[i][j] += basis[i] * basis[j];
                            atb[i] += basis[i] * target;
                        }
                    }
            }
            double a[3][4];
            memset(a, 0, sizeof a);
            for (int i = 0; i < nvar; i++) {
                for (int j = 0; j < nvar; j++) a[i][j] = ata[first_coeff + i][first_coeff + j];
                a[i][nvar] = atb[first_coeff + i];
            }
            int ok = 1;
            for (int i = 0; i < nvar; i++) {
                int piv = i;
                for (int r = i + 1; r < nvar; r++)
                    if (fabs(a[r][i]) > fabs(a[piv][i])) piv = r;
                double tmp[4];
                memcpy(tmp, a[i], sizeof tmp);
                memcpy(a[i], a[piv], sizeof tmp);
                memcpy(a[piv], tmp, sizeof tmp);
                if (fabs(a[i][i]) < 1e-9) {
                    ok = 0;
                    break;
                }
                for (int r = i + 1; r < nvar; r++) {
                    double f = a[r][i] / a[i][i];
                    for (int cc = i; cc <= nvar; cc++) a[r][cc] -= f * a[i][cc];
                }
            }
            double *out = dir == 0 ? ht : vt;
            if (ok) {
                double sol[3] = {0, 0, 0};
                for (int i = nvar - 1; i >= 0; i--) {
                    double v = a[i][nvar];
                    for (int j = i + 1; j < nvar; j++) v -= a[i][j] * sol[j];
                    sol[i] = v / a[i][i];
                }
                for (int i = first_coeff; i < 3; i++) {
                    double s = sol[i - first_coeff];
                    double lo = WIENER_TAPS_MIN[i], hi = WIENER_TAPS_MAX[i];
                    if (s < lo) s = lo;
                    if (s > hi) s = hi;
                    out[i] = s;
                }
            }
            if (trace) memcpy(trace[step], out, sizeof(double) * 3);
            step++;
        }
    }
    for (int pass = 0; pass < 2; pass++) {
        const double *t = pass == 0 ? vt : ht;
        for (int i = 0; i < 3; i++) {
            double r = round(t[i]);
            int32_t q = r != r ? 0 : (r >= 2147483647.0 ? 2147483647 : (r <= -2147483648.0 ? (-2147483647 - 1) : (int32_t)r));
            q = q < WIENER_TAPS_MIN[i] ? WIENER_TAPS_MIN[i] : (q > WIENER_TAPS_MAX[i] ? WIENER_TAPS_MAX[i] : q);
            taps[pass][i] = q;
        }
        if (first_coeff == 1) taps[pass][0] = 0;
    }
#undef AT
}

void oo_lr_fit_wiener(const int32_t *cdef, const uint16_t *src, uint32_t plane_w, uint32_t plane_h,
                      const uint32_t *rects, uint32_t nrects, int first_coeff, int32_t taps[2][3],
                      double trace[4][3]) {
    fit_wiener(cdef, src, plane_w, plane_h, rects, nrects, first_coeff, taps, trace);
}

/* f64 -> i32 as Rust's `as` casts it (saturating, NaN -> 0). */
static int32_t sat_i32(double r) {
    if (r != r) return 0;
    if (r >= 2147483647.0) return 2147483647;
    if (r <= -2147483648.0) return -2147483647 - 1;
    return (int32_t)r;
}

static int32_t clampi(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* read_delta (lr_elect.rs:580-596) into out (rect order) */
static void read_delta(const oo_ec *E, uint32_t plane, const oo_unit_blocks *ub, int32_t *out) {
    size_t stride = E->pw[plane], i = 0;
    for (uint32_t r = 0; r < ub->n; r++) {
        const uint32_t *q = ub->rects + 4 * r;
        for (uint32_t row = q[1]; row < q[1] + q[3]; row++)
            for (uint32_t col = q[0]; col < q[0] + q[2]; col++)
                out[i++] = E->lr[plane].samples[row * stride + col] - E->cdef[plane].samples[row * stride + col];
    }
}

static size_t ub_pixels(const oo_unit_blocks *ub) {
    size_t n = 0;
    for (uint32_t r = 0; r < ub->n; r++) n += (size_t)ub->rects[4 * r + 2] * ub->rects[4 * r + 3];
    return n;
}

uint32_t oo_lr_unit_rects(const oxav1_lr_elect_input *in, uint8_t plane, uint32_t ur, uint32_t uc,
                          uint32_t *rects, uint32_t cap) {
    oo_ec E;
    ec_init(&E, in, in->cdef);
    uint32_t n = 0;
    if (plane < E.np && ur < E.urows[plane] && uc < E.ucols[plane]) {
        const oo_unit_blocks *ub = &E.units[plane][ur * E.ucols[plane] + uc];
        n = ub->n;
        for (uint32_t i = 0; i < n && i < cap; i++) memcpy(rects + 4 * i, ub->rects + 4 * i, 4 * sizeof(uint32_t));
    }
    ec_free(&E);
    return n;
}

static const oxav1_lr_unit UNIT_NONE = {0, 0, {{0, 0, 0}, {0, 0, 0}}, {0, 0}};

/* elect_lr (lr_elect.rs:421-811) */
int oo_lr_elect(const oxav1_lr_elect_input *in, oxav1_lr_election *out) {
    oo_ec E;
    ec_init(&E, in, in->cdef);
    uint8_t frt[3] = {0, 0, 0};
    uint64_t d_total = 0, d_pre_total = 0;
    oo_lr_wstate st;
    wstate_init(&st);
    for (uint32_t p = 0; p < 3; p++) {
        out->unit_rows[p] = p < E.np ? E.urows[p] : 0;
        out->unit_cols[p] = p < E.np ? E.ucols[p] : 0;
    }
    for (uint32_t plane = 0; plane < E.np; plane++) {
        uint32_t urows = E.urows[plane], ucols = E.ucols[plane];
        uint32_t plane_w = E.pw[plane], plane_h = E.ph[plane];
        int first_coeff = plane != 0;
        int any_w = 0, any_s = 0;
        uint64_t plane_d = 0;
        for (uint32_t ur = 0; ur < urows; ur++)
            for (uint32_t uc = 0; uc < ucols; uc++) {
                uint32_t k = ur * ucols + uc;
                const oo_unit_blocks *ub = &E.units[plane][k];
                oxav1_lr_unit_cands cand;
                memset(&cand, 0, sizeof cand);
                if (ub->n == 0) {
                    if (out->units[plane]) out->units[plane][k] = UNIT_NONE;
                    if (out->cands[plane]) out->cands[plane][k] = cand;
                    continue;
                }
                uint64_t d_none = eval_unit(&E, plane, ur, uc, &UNIT_NONE, ub);
                d_pre_total += d_none;
                cand.d_none = d_none;

                /* Wiener candidate */
                int32_t taps[2][3];
                fit_wiener(E.cdef[plane].samples, in->input[plane], plane_w, plane_h, ub->rects, ub->n,
                           first_coeff, taps, NULL);
                oxav1_lr_unit wu = UNIT_NONE;
                wu.type = OXAV1_RESTORE_WIENER;
                for (int a = 0; a < 2; a++)
                    for (int b = 0; b < 3; b++) {
                        wu.wiener[a][b] = (int16_t)taps[a][b];
                        cand.wiener[a][b] = (int16_t)taps[a][b];
                    }
                uint64_t d_wiener = eval_unit(&E, plane, ur, uc, &wu, ub);
                cand.d_wiener = d_wiener;

                /* self-guided: probe-fit each set, keep the best by SSD */
                size_t np_ = ub_pixels(ub);
                int32_t *a0 = (int32_t *)malloc(sizeof(int32_t) * np_);
                int32_t *a1 = (int32_t *)malloc(sizeof(int32_t) * np_);
                double *t = (double *)malloc(sizeof(double) * np_);
                int have_best = 0;
                oxav1_lr_unit best_sgr = UNIT_NONE;
                uint64_t best_sgr_d = 0;
                for (uint32_t set = 0; set < 16; set++) {
                    int32_t r0 = oo_sgr_params[set * 4 + 0], r1 = oo_sgr_params[set * 4 + 2];
                    oxav1_lr_unit pu = UNIT_NONE;
                    pu.type = OXAV1_RESTORE_SGRPROJ;
                    pu.sgr_set = (uint8_t)set;
                    if (r0 != 0) {
                        pu.sgr_xqd[0] = 128;
                        pu.sgr_xqd[1] = 0;
                        (void)eval_unit(&E, plane, ur, uc, &pu, ub);
                        read_delta(&E, plane, ub, a0);
                    }
                    if (r1 != 0) {
                        pu.sgr_xqd[0] = 0;
                        pu.sgr_xqd[1] = 128;
                        (void)eval_unit(&E, plane, ur, uc, &pu, ub);
                        read_delta(&E, plane, ub, a1);
                    }
                    {
                        size_t i = 0;
                        size_t stride = plane_w;
                        for (uint32_t r = 0; r < ub->n; r++) {
                            const uint32_t *q = ub->rects + 4 * r;
                            for (uint32_t row = q[1]; row < q[1] + q[3]; row++)
                                for (uint32_t col = q[0]; col < q[0] + q[2]; col++)
                                    t[i++] = 128.0 * ((double)in->input[plane][row * stride + col] -
                                                      (double)E.cdef[plane].samples[row * stride + col]);
                        }
                    }
                    int32_t xq0 = 0, xq1 = 0;
                    if (r0 != 0 && r1 != 0) {
                        double s00 = 0, s01 = 0, s11 = 0, b0 = 0, b1 = 0;
                        for (size_t i = 0; i < np_; i++) {
                            double x0 = (double)a0[i], x1 = (double)a1[i];
                            s00 += x0 * x0;
                            s01 += x0 * x1;
                            s11 += x1 * x1;
                            b0 += x0 * t[i];
                            b1 += x1 * t[i];
                        }
                        double det = s00 * s11 - s01 * s01;
                        if (fabs(det) > 1e-6) {
                            xq0 = sat_i32(round((s11 * b0 - s01 * b1) / det));
                            xq1 = sat_i32(round((s00 * b1 - s01 * b0) / det));
                        }
                    } else if (r0 != 0) {
                        double s00 = 0, b0 = 0;
                        for (size_t i = 0; i < np_; i++) {
                            double x0 = (double)a0[i];
                            s00 += x0 * x0;
                            b0 += x0 * t[i];
                        }
                        if (s00 > 1e-6) xq0 = sat_i32(round(b0 / s00));
                    } else if (r1 != 0) {
                        double s11 = 0, b1 = 0;
                        for (size_t i = 0; i < np_; i++) {
                            double x1 = (double)a1[i];
                            s11 += x1 * x1;
                            b1 += x1 * t[i];
                        }
                        if (s11 > 1e-6) xq1 = sat_i32(round(b1 / s11));
                    }
                    xq0 = clampi(xq0, SGRPROJ_XQD_MIN[0], SGRPROJ_XQD_MAX[0]);
                    xq1 = clampi(xq1, SGRPROJ_XQD_MIN[1], SGRPROJ_XQD_MAX[1]);
                    if (r0 == 0) xq0 = 0;
                    if (r1 == 0) xq1 = clampi(128 - xq0, SGRPROJ_XQD_MIN[1], SGRPROJ_XQD_MAX[1]);
                    oxav1_lr_unit cu = UNIT_NONE;
                    cu.type = OXAV1_RESTORE_SGRPROJ;
                    cu.sgr_set = (uint8_t)set;
                    cu.sgr_xqd[0] = (int16_t)xq0;
                    cu.sgr_xqd[1] = (int16_t)xq1;
                    uint64_t d = eval_unit(&E, plane, ur, uc, &cu, ub);
                    cand.d_sgr[set] = d;
                    cand.sgr_xqd[set][0] = (int16_t)xq0;
                    cand.sgr_xqd[set][1] = (int16_t)xq1;
                    if (!have_best || d < best_sgr_d) {
                        best_sgr = cu;
                        best_sgr_d = d;
                        have_best = 1;
                    }
                }
                free(a0);
                free(a1);
                free(t);

                /* argmin D + lambda * R */
                uint64_t r_none = price_unit(&st, (int)plane, &UNIT_NONE, in->restoration_type_cdf);
                uint64_t r_wiener = price_unit(&st, (int)plane, &wu, in->restoration_type_cdf);
                oxav1_lr_unit best = UNIT_NONE;
                uint64_t best_d = d_none, best_s = d_none * 256 + in->lambda * r_none;
                uint64_t s_wiener = d_wiener * 256 + in->lambda * r_wiener;
                if (s_wiener < best_s) {
                    best = wu;
                    best_d = d_wiener;
                    best_s = s_wiener;
                }
                if (have_best) {
                    uint64_t r_sgr = price_unit(&st, (int)plane, &best_sgr, in->restoration_type_cdf);
                    uint64_t s_sgr = best_sgr_d * 256 + in->lambda * r_sgr;
                    if (s_sgr < best_s) {
                        best = best_sgr;
                        best_d = best_sgr_d;
                        best_s = s_sgr;
                    }
                }
                {
                    oo_cw w;
                    cw_init(&w, 0x8000);
                    (void)write_lr_unit(&w, &st, (int)plane, &best, in->restoration_type_cdf);
                }
                if (best.type == OXAV1_RESTORE_WIENER) any_w = 1;
                if (best.type == OXAV1_RESTORE_SGRPROJ) any_s = 1;
                plane_d += best_d;
                if (out->units[plane]) out->units[plane][k] = best;
                if (out->cands[plane]) out->cands[plane][k] = cand;
            }
        frt[plane] = !any_w && !any_s ? OXAV1_RESTORE_NONE
                     : (any_w && !any_s) ? OXAV1_RESTORE_WIENER
                     : (!any_w && any_s) ? OXAV1_RESTORE_SGRPROJ
                                         : OXAV1_RESTORE_SWITCHABLE;
        d_total += plane_d;
    }
    ec_free(&E);
    int uses = frt[0] != 0 || frt[1] != 0 || frt[2] != 0;
    out->elected = uses;
    for (int p = 0; p < 3; p++) {
        out->frame_restoration_type[p] = frt[p];
        out->loop_restoration_size[p] = uses ? 64 : 0;
    }
    out->uses_lr = (uint8_t)uses;
    out->uses_chroma_lr = (uint8_t)(frt[1] != 0 || frt[2] != 0);
    out->d = d_total;
    out->d_pre = d_pre_total;
    return 0;
}

/* apply_lr_plan (lr_elect.rs:829-912) */
int oo_lr_apply_plan(const oxav1_lr_election *plan, const oxav1_lr_elect_input *in, uint16_t *const recon[3],
                     uint64_t *ssd) {
    oo_ec E;
    ec_init(&E, in, (const uint16_t *const *)recon);
    oxav1_lr_ctx ctx = E.ctx;
    for (int p = 0; p < 3; p++) {
        ctx.frame_restoration_type[p] = plan->frame_restoration_type[p];
        ctx.loop_restoration_size[p] = plan->loop_restoration_size[p];
    }
    ctx.uses_lr = plan->uses_lr;
    oxav1_lr_grids g;
    memset(&g, 0, sizeof g);
    for (uint32_t p = 0; p < E.np; p++)
        if (plan->frame_restoration_type[p] != OXAV1_RESTORE_NONE) g.units[p] = plan->units[p];
    oo_loop_restoration_frame(&ctx, &g, E.curr, E.np, E.cdef, E.np, E.lr, E.np);
    uint64_t s = 0;
    for (uint32_t p = 0; p < E.np; p++) {
        size_t n = (size_t)E.pw[p] * E.ph[p];
        for (size_t i = 0; i < n; i++) {
            int32_t o = E.lr[p].samples[i];
            recon[p][i] = (uint16_t)(o > 0 ? o : 0);
            int64_t d = (int64_t)o - (int64_t)in->input[p][i];
            s += (uint64_t)(d * d);
        }
    }
    ec_free(&E);
    *ssd = s;
    return 0;
}
