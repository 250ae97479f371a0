/*
 * oo_elect.c — CPU oracle of the encoder's CDEF strength election
 * (encoder/cdef_elect.rs).  TEST INFRASTRUCTURE ONLY (see oxav1_oracle.h).
 *
 * A literal restatement: every candidate is one full §7.15 run of the
 * oracle's cdef_frame over the reconstruction (all planes, as the reference
 * runs it), followed by the per-64x64-unit SSD against the source, in the
 * order and with the tie-breaking of the reference's search.
 *
 * Parity pins (tests/test_oracle_elect.py): the per-unit SSD tables against
 * a numpy sum over the oracle's cdef_frame output; every returned plan's D
 * against the SSD of applying the plan through cdef_frame with the plan's
 * unit grid; the winner's score against the unfiltered and frame-level
 * arms; the frame-level arm against a brute-force minimum over the coarse
 * candidate grid; per-unit choices as argmins of the tables; and the
 * reference's ringing tripwire (tests/cdef_ab.rs: edge content must elect
 * the filter and gain).
 */
#include "oxav1_oracle.h"

#include <stdlib.h>
#include <string.h>

typedef struct {
    uint8_t pri, sec;
} oo_str;

static int str_eq(oo_str a, oo_str b) { return a.pri == b.pri && a.sec == b.sec; }
static const oo_str ZERO_STR = {0, 0};

/* params_for (cdef_elect.rs:180-190) */
static oxav1_cdef_params params_for(uint8_t damping, oo_str y, oo_str uv) {
    oxav1_cdef_params p;
    memset(&p, 0, sizeof p);
    p.cdef_damping = damping;
    p.cdef_bits = 0;
    p.y_pri_strength[0] = y.pri;
    p.y_sec_strength[0] = y.sec;
    p.uv_pri_strength[0] = uv.pri;
    p.uv_sec_strength[0] = uv.sec;
    return p;
}

typedef struct {
    const oxav1_cdef_elect_input *in;
    uint32_t n_units, sb_cols, np;
    uint32_t pw[3], ph[3];
    int32_t *src[3], *dst[3];
} oo_run;

/* run_units (cdef_elect.rs:192-240): one cdef_frame run under `params`
 * over the committed grid, then the per-unit SSD of plane set `which`. */
static void run_units(oo_run *R, const oxav1_cdef_params *params, int which, uint64_t *ssd) {
    const oxav1_cdef_elect_input *in = R->in;
    oxav1_cdef_ctx ctx;
    memset(&ctx, 0, sizeof ctx);
    ctx.mi_rows = in->mi_rows;
    ctx.mi_cols = in->mi_cols;
    ctx.num_planes = in->num_planes;
    ctx.bit_depth = in->bit_depth;
    ctx.subsampling_x = in->subsampling_x;
    ctx.subsampling_y = in->subsampling_y;
    ctx.params = *params;
    oxav1_cdef_grids g = {in->cdef_idx, in->skip};
    oxav1_plane_i32 s[3], d[3];
    for (uint32_t p = 0; p < R->np; p++) {
        s[p].samples = R->src[p];
        s[p].rows = R->ph[p];
        s[p].cols = R->pw[p];
        d[p].samples = R->dst[p];
        d[p].rows = R->ph[p];
        d[p].cols = R->pw[p];
    }
    oo_cdef_frame(&ctx, &g, s, R->np, d, R->np);
    memset(ssd, 0, sizeof(uint64_t) * R->n_units);
    if (which == 0) {
        for (size_t i = 0; i < (size_t)in->width * in->height; i++) {
            size_t y = i / in->width, x = i % in->width;
            int64_t df = (int64_t)R->dst[0][i] - (int64_t)in->input[0][i];
            ssd[(y / 64) * R->sb_cols + x / 64] += (uint64_t)(df * df);
        }
    } else {
        uint32_t ssx = in->subsampling_x, ssy = in->subsampling_y;
        for (int p = 1; p < 3; p++)
            for (size_t i = 0; i < (size_t)in->chroma_w * in->chroma_h; i++) {
                size_t cy = i / in->chroma_w, cx = i % in->chroma_w;
                int64_t df = (int64_t)R->dst[p][i] - (int64_t)in->input[p][i];
                ssd[((cy << ssy) / 64) * R->sb_cols + (cx << ssx) / 64] += (uint64_t)(df * df);
            }
    }
}

/* SetTables (cdef_elect.rs:115-134): cands[0] is the zero strength. */
typedef struct {
    int n;
    oo_str cands[64];
    uint64_t *ssd[64];
} oo_tab;

static uint64_t tab_total(const oo_tab *t, uint32_t n_units, int c) {
    uint64_t s = 0;
    for (uint32_t k = 0; k < n_units; k++) s += t->ssd[c][k];
    return s;
}
/* min_by_key: the first minimum */
static int tab_best(const oo_tab *t, uint32_t n_units) {
    int best = 0;
    uint64_t bv = 0;
    for (int c = 0; c < t->n; c++) {
        uint64_t v = tab_total(t, n_units, c);
        if (c == 0 || v < bv) {
            bv = v;
            best = c;
        }
    }
    return best;
}
static int tab_has(const oo_tab *t, oo_str s) {
    for (int c = 0; c < t->n; c++)
        if (str_eq(t->cands[c], s)) return 1;
    return 0;
}

/* the sweep's eval closure (cdef_elect.rs:276-288) */
static void tab_eval(oo_run *R, oo_tab *t, uint8_t damping, int which, oo_str s) {
    if (tab_has(t, s)) return;
    oxav1_cdef_params p = which == 0 ? params_for(damping, s, ZERO_STR) : params_for(damping, ZERO_STR, s);
    t->ssd[t->n] = (uint64_t *)malloc(sizeof(uint64_t) * R->n_units);
    run_units(R, &p, which, t->ssd[t->n]);
    t->cands[t->n] = s;
    t->n++;
}

/* sweep (cdef_elect.rs:270-305): coarse grid, secondary-only, refine */
static void sweep(oo_run *R, oo_tab *t, uint8_t damping, int which, const uint64_t *base) {
    t->n = 1;
    t->cands[0] = ZERO_STR;
    t->ssd[0] = (uint64_t *)malloc(sizeof(uint64_t) * R->n_units);
    memcpy(t->ssd[0], base, sizeof(uint64_t) * R->n_units);
    static const uint8_t pris[8] = {1, 2, 3, 4, 6, 9, 12, 15};
    for (int i = 0; i < 8; i++)
        for (int j = 0; j < 2; j++) tab_eval(R, t, damping, which, (oo_str){pris[i], (uint8_t)(j ? 2 : 0)});
    static const uint8_t secs[3] = {1, 2, 4};
    for (int j = 0; j < 3; j++) tab_eval(R, t, damping, which, (oo_str){0, secs[j]});
    oo_str center = t->cands[tab_best(t, R->n_units)];
    int lo = center.pri >= 1 ? center.pri - 1 : 0;
    int hi = center.pri + 1 < 15 ? center.pri + 1 : 15;
    static const uint8_t rsec[4] = {0, 1, 2, 4};
    for (int pri = lo; pri <= hi; pri++)
        for (int j = 0; j < 4; j++) tab_eval(R, t, damping, which, (oo_str){(uint8_t)pri, rsec[j]});
}

static void tab_free(oo_tab *t) {
    for (int c = 0; c < t->n; c++) free(t->ssd[c]);
    t->n = 0;
}

typedef struct {
    uint64_t d, r256;
    uint8_t damping, bits;
    int table; /* -1 base, -2 frame level, else damping index */
    int nsets;
    int sets[8][2];
    int *choice;
} oo_plan;

static uint64_t score256(uint64_t d, uint64_t lambda, uint64_t r256) { return d * 256 + lambda * r256; }

/* elect_cdef (cdef_elect.rs:136-575) */
int oo_cdef_elect(const oxav1_cdef_elect_input *in, oxav1_cdef_election *out) {
    int8_t *best_ids = out->best_unit_idx, *fl_ids = out->frame_level_unit_idx;
    memset(out, 0, sizeof *out);
    out->best_unit_idx = best_ids;
    out->frame_level_unit_idx = fl_ids;
    const uint32_t sb_rows = (in->mi_rows + 15) / 16, sb_cols = (in->mi_cols + 15) / 16;
    const uint32_t n_units = sb_rows * sb_cols;
    /* coded units: the committed anchor is not -1 (cdef_elect.rs:143-155) */
    uint8_t *coded = (uint8_t *)calloc(n_units ? n_units : 1, 1);
    uint64_t coded_count = 0;
    for (uint32_t k = 0; k < n_units; k++) {
        coded[k] = in->cdef_idx[k] != -1;
        coded_count += coded[k];
    }
    if (coded_count == 0) {
        free(coded);
        return OXAV1_OK;
    }
    oo_run R;
    memset(&R, 0, sizeof R);
    R.in = in;
    R.n_units = n_units;
    R.sb_cols = sb_cols;
    R.np = in->num_planes > 1 ? 3 : 1;
    for (uint32_t p = 0; p < R.np; p++) {
        R.pw[p] = p ? in->chroma_w : in->width;
        R.ph[p] = p ? in->chroma_h : in->height;
        size_t n = (size_t)R.pw[p] * R.ph[p];
        R.src[p] = (int32_t *)malloc(sizeof(int32_t) * (n ? n : 1));
        R.dst[p] = (int32_t *)malloc(sizeof(int32_t) * (n ? n : 1));
        for (size_t i = 0; i < n; i++) R.src[p][i] = in->recon[p][i];
        memcpy(R.dst[p], R.src[p], sizeof(int32_t) * n);
    }
    /* unfiltered baselines (cdef_elect.rs:243-268) */
    uint64_t *base_y = (uint64_t *)calloc(n_units, sizeof(uint64_t));
    uint64_t *base_uv = (uint64_t *)calloc(n_units, sizeof(uint64_t));
    for (size_t i = 0; i < (size_t)in->width * in->height; i++) {
        size_t y = i / in->width, x = i % in->width;
        int64_t df = (int64_t)in->recon[0][i] - (int64_t)in->input[0][i];
        base_y[(y / 64) * sb_cols + x / 64] += (uint64_t)(df * df);
    }
    if (in->num_planes > 1)
        for (int p = 1; p < 3; p++)
            for (size_t i = 0; i < (size_t)in->chroma_w * in->chroma_h; i++) {
                size_t cy = i / in->chroma_w, cx = i % in->chroma_w;
                int64_t df = (int64_t)in->recon[p][i] - (int64_t)in->input[p][i];
                base_uv[((cy << in->subsampling_y) / 64) * sb_cols + (cx << in->subsampling_x) / 64] +=
                    (uint64_t)(df * df);
            }
    uint64_t base_total = 0;
    for (uint32_t k = 0; k < n_units; k++) base_total += base_y[k] + base_uv[k];

    /* per-damping tables (cdef_elect.rs:307-322) */
    static const uint8_t dampings[2] = {3, 5};
    oo_tab ty[2], tuv[2];
    for (int di = 0; di < 2; di++) {
        sweep(&R, &ty[di], dampings[di], 0, base_y);
        if (in->num_planes > 1) {
            sweep(&R, &tuv[di], dampings[di], 1, base_uv);
        } else {
            tuv[di].n = 1;
            tuv[di].cands[0] = ZERO_STR;
            tuv[di].ssd[0] = (uint64_t *)malloc(sizeof(uint64_t) * n_units);
            memcpy(tuv[di].ssd[0], base_uv, sizeof(uint64_t) * n_units);
        }
    }
    /* frame-level arm (cdef_elect.rs:324-356) */
    uint8_t fl_damping = 3;
    oo_str fl_y = ZERO_STR, fl_uv = ZERO_STR;
    uint64_t fl_total = UINT64_MAX;
    for (int di = 0; di < 2; di++) {
        int yi = tab_best(&ty[di], n_units), uvi = tab_best(&tuv[di], n_units);
        uint64_t total = tab_total(&ty[di], n_units, yi) + tab_total(&tuv[di], n_units, uvi);
        if (total < fl_total) {
            fl_damping = dampings[di];
            fl_y = ty[di].cands[yi];
            fl_uv = tuv[di].cands[uvi];
            fl_total = total;
        }
    }
    uint64_t *tmp = (uint64_t *)malloc(sizeof(uint64_t) * n_units);
    if (!str_eq(fl_y, ZERO_STR) || !str_eq(fl_uv, ZERO_STR)) {
        static const uint8_t refine[2] = {4, 6};
        for (int i = 0; i < 2; i++) {
            oxav1_cdef_params p = params_for(refine[i], fl_y, fl_uv);
            uint64_t t = 0;
            run_units(&R, &p, 0, tmp);
            for (uint32_t k = 0; k < n_units; k++) t += tmp[k];
            if (in->num_planes > 1) {
                run_units(&R, &p, 1, tmp);
                for (uint32_t k = 0; k < n_units; k++) t += tmp[k];
            }
            if (t < fl_total) {
                fl_total = t;
                fl_damping = refine[i];
            }
        }
    }
    free(tmp);

    /* plans (cdef_elect.rs:358-466) */
    oo_plan plans[16];
    int nplans = 0;
    memset(plans, 0, sizeof plans);
    plans[nplans].d = base_total;
    plans[nplans].damping = 3;
    plans[nplans].table = -1;
    nplans++;
    if (fl_total < UINT64_MAX && (!str_eq(fl_y, ZERO_STR) || !str_eq(fl_uv, ZERO_STR))) {
        plans[nplans].d = fl_total;
        plans[nplans].damping = fl_damping;
        plans[nplans].table = -2;
        nplans++;
    }
    const uint64_t set_hdr_bits = in->num_planes > 1 ? 12 : 6;
    uint64_t *cur = (uint64_t *)malloc(sizeof(uint64_t) * n_units);
    for (int di = 0; di < 2 && in->max_bits > 0; di++) {
        const oo_tab *a = &ty[di], *b = &tuv[di];
#define UNIT_COST(yi, uvi, k) (a->ssd[(yi)][(k)] + b->ssd[(uvi)][(k)])
        int best_single[2] = {0, 0};
        uint64_t best_single_d = UINT64_MAX;
        for (int yi = 0; yi < a->n; yi++)
            for (int uvi = 0; uvi < b->n; uvi++) {
                uint64_t t = tab_total(a, n_units, yi) + tab_total(b, n_units, uvi);
                if (t < best_single_d) {
                    best_single_d = t;
                    best_single[0] = yi;
                    best_single[1] = uvi;
                }
            }
        int sets[8][2];
        int nsets = 1;
        sets[0][0] = best_single[0];
        sets[0][1] = best_single[1];
        for (uint32_t k = 0; k < n_units; k++) cur[k] = UNIT_COST(best_single[0], best_single[1], k);
        int maxb = in->max_bits < 3 ? in->max_bits : 3;
        for (int bits = 1; bits <= maxb; bits++) {
            int want = 1 << bits;
            while (nsets < want) {
                uint64_t best_gain = 0;
                int found = 0, bs[2] = {0, 0};
                for (int yi = 0; yi < a->n; yi++)
                    for (int uvi = 0; uvi < b->n; uvi++) {
                        int dup = 0;
                        for (int s = 0; s < nsets; s++)
                            if (sets[s][0] == yi && sets[s][1] == uvi) dup = 1;
                        if (dup) continue;
                        uint64_t gain = 0;
                        for (uint32_t k = 0; k < n_units; k++) {
                            if (!coded[k]) continue;
                            uint64_t c = UNIT_COST(yi, uvi, k);
                            gain += cur[k] > c ? cur[k] - c : 0;
                        }
                        if (gain > best_gain) {
                            best_gain = gain;
                            found = 1;
                            bs[0] = yi;
                            bs[1] = uvi;
                        }
                    }
                if (found) {
                    for (uint32_t k = 0; k < n_units; k++)
                        if (coded[k]) {
                            uint64_t c = UNIT_COST(bs[0], bs[1], k);
                            if (c < cur[k]) cur[k] = c;
                        }
                    sets[nsets][0] = bs[0];
                    sets[nsets][1] = bs[1];
                } else {
                    sets[nsets][0] = best_single[0];
                    sets[nsets][1] = best_single[1];
                }
                nsets++;
            }
            oo_plan *P = &plans[nplans++];
            P->choice = (int *)malloc(sizeof(int) * n_units);
            uint64_t dtot = 0;
            for (uint32_t k = 0; k < n_units; k++) {
                int ch = 0;
                if (coded[k]) {
                    uint64_t bv = UNIT_COST(sets[0][0], sets[0][1], k);
                    for (int s = 1; s < nsets; s++) {
                        uint64_t v = UNIT_COST(sets[s][0], sets[s][1], k);
                        if (v < bv) {
                            bv = v;
                            ch = s;
                        }
                    }
                    dtot += bv;
                } else {
                    dtot += base_y[k] + base_uv[k];
                }
                P->choice[k] = ch;
            }
            P->d = dtot;
            P->r256 = 256 * ((uint64_t)bits * coded_count + ((1ull << bits) - 1) * set_hdr_bits);
            P->damping = dampings[di];
            P->bits = (uint8_t)bits;
            P->table = di;
            P->nsets = nsets;
            memcpy(P->sets, sets, sizeof sets);
        }
#undef UNIT_COST
    }
    free(cur);

    /* winner: the first minimum of D*256 + lambda*R256 (cdef_elect.rs:520-523) */
    int w = 0;
    for (int i = 1; i < nplans; i++)
        if (score256(plans[i].d, in->lambda, plans[i].r256) < score256(plans[w].d, in->lambda, plans[w].r256)) w = i;
    out->base_d = base_total;
    const int fl_present = (!str_eq(fl_y, ZERO_STR) || !str_eq(fl_uv, ZERO_STR)) && fl_total < base_total;
    if (plans[w].table == -1) {
        out->elected = 0;
    } else if (plans[w].table == -2) {
        /* frame-level arm won outright (cdef_elect.rs:538-546) */
        out->elected = 1;
        out->best = params_for(fl_damping, fl_y, fl_uv);
        out->best_d = fl_total;
        if (best_ids)
            for (uint32_t k = 0; k < n_units; k++) best_ids[k] = coded[k] ? 0 : -1;
        out->has_frame_level = 0;
    } else {
        /* per-unit arm (cdef_elect.rs:547-574) */
        const oo_plan *P = &plans[w];
        const oo_tab *a = &ty[P->table], *b = &tuv[P->table];
        oxav1_cdef_params p;
        memset(&p, 0, sizeof p);
        p.cdef_damping = P->damping;
        p.cdef_bits = P->bits;
        for (int i = 0; i < P->nsets; i++) {
            p.y_pri_strength[i] = a->cands[P->sets[i][0]].pri;
            p.y_sec_strength[i] = a->cands[P->sets[i][0]].sec;
            p.uv_pri_strength[i] = b->cands[P->sets[i][1]].pri;
            p.uv_sec_strength[i] = b->cands[P->sets[i][1]].sec;
        }
        out->elected = 1;
        out->best = p;
        out->best_d = P->d;
        if (best_ids)
            for (uint32_t k = 0; k < n_units; k++) best_ids[k] = coded[k] ? (int8_t)P->choice[k] : -1;
        out->has_frame_level = fl_present;
        if (fl_present) {
            out->frame_level = params_for(fl_damping, fl_y, fl_uv);
            out->frame_level_d = fl_total;
            if (fl_ids)
                for (uint32_t k = 0; k < n_units; k++) fl_ids[k] = coded[k] ? 0 : -1;
        }
    }
    for (int i = 0; i < nplans; i++) free(plans[i].choice);
    for (int di = 0; di < 2; di++) {
        tab_free(&ty[di]);
        tab_free(&tuv[di]);
    }
    for (uint32_t p = 0; p < R.np; p++) {
        free(R.src[p]);
        free(R.dst[p]);
    }
    free(base_y);
    free(base_uv);
    free(coded);
    return OXAV1_OK;
}

/* Test helper: run_units for one candidate parameter set (the SSD tables
 * the search reads), plane set `which`. */
void oo_cdef_run_units(const oxav1_cdef_elect_input *in, const oxav1_cdef_params *params, int which, uint64_t *ssd) {
    oo_run R;
    memset(&R, 0, sizeof R);
    R.in = in;
    R.sb_cols = (in->mi_cols + 15) / 16;
    R.n_units = ((in->mi_rows + 15) / 16) * R.sb_cols;
    R.np = in->num_planes > 1 ? 3 : 1;
    for (uint32_t p = 0; p < R.np; p++) {
        R.pw[p] = p ? in->chroma_w : in->width;
        R.ph[p] = p ? in->chroma_h : in->height;
        size_t n = (size_t)R.pw[p] * R.ph[p];
        R.src[p] = (int32_t *)malloc(sizeof(int32_t) * (n ? n : 1));
        R.dst[p] = (int32_t *)malloc(sizeof(int32_t) * (n ? n : 1));
        for (size_t i = 0; i < n; i++) R.src[p][i] = in->recon[p][i];
        memcpy(R.dst[p], R.src[p], sizeof(int32_t) * n);
    }
    run_units(&R, params, which, ssd);
    for (uint32_t p = 0; p < R.np; p++) {
        free(R.src[p]);
        free(R.dst[p]);
    }
}

/* ===================================================================== */
/* Superres election helpers (encoder/superres_elect.rs)                  */
/* ===================================================================== */

/* superres_coded_width (superres_elect.rs:23-29): SUPERRES_NUM = 8 */
uint32_t oo_superres_coded_width(uint32_t upscaled_width, uint32_t denom) {
    return (upscaled_width * 8 + denom / 2) / denom;
}

/* candidate_denoms (superres_elect.rs:31-45): SUPERRES_DENOM_MIN = 9 */
uint32_t oo_superres_candidate_denoms(uint32_t width, uint32_t *out) {
    uint32_t n = 0;
    for (uint32_t d = 9; d <= 16; d++) {
        uint32_t wd = oo_superres_coded_width(width, d);
        if (wd >= 8 && wd % 8 == 0 && wd < width) out[n++] = d;
    }
    return n;
}

/* superres_probe (superres_elect.rs:62-91) */
int oo_superres_probe(const uint16_t *y, uint32_t w, uint32_t h, uint8_t bit_depth) {
    if (w < 3) return 0;
    uint64_t sum = 0, count = 0;
    for (size_t r = 0; r < h; r++)
        for (size_t c = 1; c < (size_t)w - 1; c++) {
            int64_t m = y[r * w + c];
            int64_t v = 2 * m - (int64_t)y[r * w + c - 1] - (int64_t)y[r * w + c + 1];
            sum += (uint64_t)(v < 0 ? -v : v);
            count++;
        }
    if (count == 0) return 0;
    uint64_t mean16 = (sum * 16) / count / (1ull << (bit_depth - 8));
    return mean16 < 10;
}

/* downscale_plane_width (superres_elect.rs:125-150) */
static void downscale_plane_width(const uint16_t *src, size_t w_src, size_t h, size_t w_dst, uint16_t *out) {
    uint32_t *lp = (uint32_t *)malloc(sizeof(uint32_t) * (w_src ? w_src : 1));
    for (size_t r = 0; r < h; r++) {
        const uint16_t *row = src + r * w_src;
        for (size_t x = 0; x < w_src; x++) {
            uint32_t l = row[x ? x - 1 : 0], m = row[x], rr = row[x + 1 < w_src ? x + 1 : w_src - 1];
            lp[x] = (l + 2 * m + rr + 2) >> 2;
        }
        for (size_t x = 0; x < w_dst; x++) {
            uint64_t s64 = ((2 * (uint64_t)x + 1) * (uint64_t)w_src * 32) / (uint64_t)w_dst - 32;
            size_t i0 = (size_t)(s64 >> 6);
            uint32_t frac = (uint32_t)(s64 & 63);
            size_t i1 = i0 + 1 < w_src ? i0 + 1 : w_src - 1;
            out[r * w_dst + x] = (uint16_t)(((64 - frac) * lp[i0] + frac * lp[i1] + 32) >> 6);
        }
    }
    free(lp);
}

/* downscale_width (superres_elect.rs:93-123): luma to coded_width, chroma
 * from width >> ssx to coded_width >> ssx columns, heights unchanged. */
void oo_superres_downscale(const uint16_t *const src[3], uint32_t width, uint32_t height, uint8_t ssx, uint8_t ssy,
                           uint8_t num_planes, uint32_t coded_width, uint16_t *const dst[3]) {
    downscale_plane_width(src[0], width, height, coded_width, dst[0]);
    if (num_planes > 1)
        for (int p = 1; p < 3; p++)
            downscale_plane_width(src[p], width >> ssx, height >> ssy, coded_width >> ssx, dst[p]);
}

/* upscale_recon (superres_elect.rs:152-232): §7.16 on u16 planes of
 * (width + ssx) >> ssx etc., narrowed back to u16. */
int oo_superres_upscale_recon(const uint16_t *const planes[3], uint32_t width, uint32_t height,
                              uint32_t upscaled_width, uint32_t mi_cols, uint8_t bit_depth, uint8_t ssx, uint8_t ssy,
                              uint8_t num_planes, uint16_t *const out[3]) {
    oxav1_sr_ctx ctx;
    memset(&ctx, 0, sizeof ctx);
    ctx.use_superres = 1;
    ctx.frame_width = width;
    ctx.upscaled_width = upscaled_width;
    ctx.frame_height = height;
    ctx.mi_cols = mi_cols;
    ctx.num_planes = num_planes;
    ctx.bit_depth = bit_depth;
    ctx.subsampling_x = ssx;
    ctx.subsampling_y = ssy;
    oxav1_plane_i32 in[3], o[3];
    memset(in, 0, sizeof in);
    memset(o, 0, sizeof o);
    for (uint32_t p = 0; p < num_planes && p < 3; p++) {
        uint32_t pw = p ? (width + ssx) >> ssx : width, uw = p ? (upscaled_width + ssx) >> ssx : upscaled_width;
        uint32_t ph = p ? (height + ssy) >> ssy : height;
        in[p].rows = ph;
        in[p].cols = pw;
        in[p].samples = (int32_t *)malloc(sizeof(int32_t) * ((size_t)pw * ph + 1));
        for (size_t i = 0; i < (size_t)pw * ph; i++) in[p].samples[i] = planes[p][i];
        o[p].rows = ph;
        o[p].cols = uw;
        o[p].samples = (int32_t *)calloc((size_t)uw * ph + 1, sizeof(int32_t));
    }
    int st = oo_upscale_frame(&ctx, in, num_planes, o, num_planes);
    for (uint32_t p = 0; p < num_planes && p < 3; p++) {
        if (!st)
            for (size_t i = 0; i < (size_t)o[p].rows * o[p].cols; i++) {
                int32_t v = o[p].samples[i];
                out[p][i] = (uint16_t)(v < 0 ? 0 : (v > 65535 ? 65535 : v));
            }
        free(in[p].samples);
        free(o[p].samples);
    }
    return st;
}

/* ===================================================================== */
/* Film-grain election helpers (encoder/film_grain_elect.rs)              */
/* ===================================================================== */
#include <math.h>

/* denoise_plane (film_grain_elect.rs:71-94): [1,2,1] x [1,2,1] / 16 with
 * edge replication */
static void denoise_plane(const uint16_t *src, uint16_t *dst, size_t w, size_t h) {
    if (w == 0 || h == 0) return;
    for (size_t r = 0; r < h; r++) {
        size_t rm = r ? r - 1 : 0, rp = r + 1 < h ? r + 1 : h - 1;
        for (size_t c = 0; c < w; c++) {
            size_t cm = c ? c - 1 : 0, cp = c + 1 < w ? c + 1 : w - 1;
#define AT(rr, cc) ((uint32_t)src[(rr) * w + (cc)])
            uint32_t sum = AT(rm, cm) + 2 * AT(rm, c) + AT(rm, cp) + 2 * AT(r, cm) + 4 * AT(r, c) + 2 * AT(r, cp) +
                           AT(rp, cm) + 2 * AT(rp, c) + AT(rp, cp);
#undef AT
            dst[r * w + c] = (uint16_t)((sum + 8) >> 4);
        }
    }
}

/* denoise_frame (film_grain_elect.rs:52-69): chroma planes are
 * (width >> ssx) x (height >> ssy) */
void oo_fg_denoise(const uint16_t *const src[3], uint32_t width, uint32_t height, uint8_t ssx, uint8_t ssy,
                   uint8_t num_planes, uint16_t *const dst[3]) {
    denoise_plane(src[0], dst[0], width, height);
    if (num_planes > 1) {
        denoise_plane(src[1], dst[1], width >> ssx, height >> ssy);
        denoise_plane(src[2], dst[2], width >> ssx, height >> ssy);
    }
}

/* sigma16 closure (film_grain_elect.rs:103-108) */
static uint64_t sigma16(uint64_t sq, uint64_t count) {
    if (count == 0) return 0;
    return (uint64_t)llround(sqrt((double)sq / (double)count) * 16.0);
}

/* chroma_profile closure (film_grain_elect.rs:130-158) */
static void chroma_profile(const uint16_t *src, const uint16_t *den, size_t n, uint32_t shift, uint64_t *g,
                           uint64_t bins[4]) {
    if (n == 0) {
        *g = 0;
        memset(bins, 0, 4 * sizeof(uint64_t));
        return;
    }
    uint64_t sq_sum = 0, bsq[4] = {0}, bn[4] = {0};
    for (size_t i = 0; i < n; i++) {
        int64_t d = ((int64_t)src[i] - (int64_t)den[i]) >> shift;
        uint64_t sq = (uint64_t)(d * d);
        sq_sum += sq;
        size_t bin = ((size_t)src[i] >> shift) >> 6;
        if (bin > 3) bin = 3;
        bsq[bin] += sq;
        bn[bin]++;
    }
    *g = sigma16(sq_sum, n);
    uint64_t minc = n / 100 > 32 ? n / 100 : 32;
    for (int i = 0; i < 4; i++) bins[i] = bn[i] >= minc ? sigma16(bsq[i], bn[i]) : *g;
}

/* noise_estimate (film_grain_elect.rs:97-167) */
void oo_fg_noise_estimate(const uint16_t *const input[3], const uint16_t *const denoised[3], uint32_t width,
                          uint32_t height, uint8_t ssx, uint8_t ssy, uint8_t num_planes, uint8_t bit_depth,
                          oxav1_noise_estimate *out) {
    uint32_t shift = bit_depth - 8u;
    size_t ny = (size_t)width * height;
    uint64_t n = ny ? ny : 1;
    uint64_t sq_sum = 0, bin_sq[8] = {0}, bin_n[8] = {0};
    for (size_t i = 0; i < ny; i++) {
        int64_t d = ((int64_t)input[0][i] - (int64_t)denoised[0][i]) >> shift;
        uint64_t sq = (uint64_t)(d * d);
        sq_sum += sq;
        size_t bin = ((size_t)input[0][i] >> shift) >> 5;
        if (bin > 7) bin = 7;
        bin_sq[bin] += sq;
        bin_n[bin]++;
    }
    memset(out, 0, sizeof *out);
    out->sigma16 = sigma16(sq_sum, n);
    uint64_t min_count = n / 100 > 64 ? n / 100 : 64;
    for (int i = 0; i < 8; i++) out->bin_sigma16[i] = bin_n[i] >= min_count ? sigma16(bin_sq[i], bin_n[i]) : out->sigma16;
    size_t nc = num_planes > 1 ? (size_t)(width >> ssx) * (height >> ssy) : 0;
    chroma_profile(num_planes > 1 ? input[1] : NULL, num_planes > 1 ? denoised[1] : NULL, nc, shift, &out->cb_sigma16,
                   out->cb_bin_sigma16);
    chroma_profile(num_planes > 1 ? input[2] : NULL, num_planes > 1 ? denoised[2] : NULL, nc, shift, &out->cr_sigma16,
                   out->cr_bin_sigma16);
}

/* plane_sigma16 (film_grain_elect.rs:896-910) */
uint64_t oo_fg_plane_sigma16(const uint16_t *a, const uint16_t *b, size_t n, uint8_t bit_depth) {
    if (n == 0) return 0;
    uint32_t shift = bit_depth - 8u;
    uint64_t sq = 0;
    for (size_t i = 0; i < n; i++) {
        int64_t d = ((int64_t)a[i] - (int64_t)b[i]) >> shift;
        sq += (uint64_t)(d * d);
    }
    return (uint64_t)llround(sqrt((double)sq / (double)n) * 16.0);
}

/* apply_grain_to_recon (film_grain_elect.rs:834-893): §7.18.3 synthesis
 * of u16 planes ((width + ssx) >> ssx etc.), narrowed with max(0) */
void oo_fg_apply_grain(const oxav1_film_grain_params *fg, const uint16_t *const planes[3], uint32_t width,
                       uint32_t height, uint8_t bit_depth, uint8_t ssx, uint8_t ssy, uint8_t num_planes,
                       uint8_t matrix_coefficients, uint16_t *const out[3]) {
    oxav1_plane_i32 b[3];
    memset(b, 0, sizeof b);
    for (uint32_t p = 0; p < num_planes && p < 3; p++) {
        uint32_t pw = p ? (width + ssx) >> ssx : width, ph = p ? (height + ssy) >> ssy : height;
        b[p].rows = ph;
        b[p].cols = pw;
        b[p].samples = (int32_t *)malloc(sizeof(int32_t) * ((size_t)pw * ph + 1));
        for (size_t i = 0; i < (size_t)pw * ph; i++) b[p].samples[i] = planes[p][i];
    }
    oo_film_grain_synthesis(fg, bit_depth, num_planes, ssx, ssy, matrix_coefficients, b, num_planes);
    for (uint32_t p = 0; p < num_planes && p < 3; p++) {
        for (size_t i = 0; i < (size_t)b[p].rows * b[p].cols; i++) {
            int32_t v = b[p].samples[i];
            out[p][i] = (uint16_t)(v < 0 ? 0 : v);
        }
        free(b[p].samples);
    }
}
