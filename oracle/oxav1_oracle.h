/*
 * oxav1_oracle.h — CPU oracle for the AV1 post-reconstruction filter chain.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the
 * reference's CPU filter path (oxideav-av1 v0.1.16: src/loop_filter.rs,
 * src/cdef.rs, src/superres.rs, src/loop_restoration.rs, src/film_grain.rs
 * and the chain in src/decoder/frame_driver.rs:790-1116).  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker or the timed CPU baseline.  The product library never
 * links or calls it.
 *
 * The closure-based contexts of the reference are taken as the same dense
 * grids the product's C ABI uses (include/oxav1_pf.h); each grid accessor
 * below mirrors the closure the reference's walker bridges build
 * (cdf.rs:20320-20678).
 *
 * Parity status of each function is recorded in oxav1_oracle.c's header.
 */
#ifndef OXAV1_ORACLE_H
#define OXAV1_ORACLE_H

#include "../include/oxav1_pf.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- §7.14 deblocking (loop_filter.rs) ---- */
typedef struct oo_lf_strength {
    uint8_t lvl, limit, blimit, thresh;
} oo_lf_strength;
typedef struct oo_lf_masks {
    uint8_t hev_mask, filter_mask, flat_mask, flat_mask2;
} oo_lf_masks;

void oo_loop_filter_frame(const oxav1_lf_ctx *ctx, const oxav1_lf_grids *g,
                          oxav1_plane_i32 *planes, uint32_t n);
/* Test helper: every edge of a pass reads the pass-start samples. */
void oo_loop_filter_frame_parallel(const oxav1_lf_ctx *ctx, const oxav1_lf_grids *g,
                                   oxav1_plane_i32 *planes, uint32_t n);
/* Test helper: the same walk with rows and columns visited in reverse. */
void oo_loop_filter_frame_reversed(const oxav1_lf_ctx *ctx, const oxav1_lf_grids *g,
                                   oxav1_plane_i32 *planes, uint32_t n);
void oo_loop_filter_edge(const oxav1_lf_ctx *ctx, const oxav1_lf_grids *g,
                         oxav1_plane_i32 *planes, uint32_t n, uint8_t plane, uint8_t pass,
                         uint32_t row, uint32_t col);
uint32_t oo_lf_filter_size(uint32_t tx_sz, uint32_t prev_tx_sz, uint8_t pass, uint8_t plane);
oo_lf_strength oo_lf_adaptive_filter_strength(const oxav1_lf_ctx *ctx, const oxav1_lf_grids *g,
                                              uint32_t row, uint32_t col, uint8_t plane,
                                              uint8_t pass);
oo_lf_masks oo_lf_filter_mask(const oxav1_plane_i32 *buf, int32_t x, int32_t y, uint8_t plane,
                              uint8_t limit, uint8_t blimit, uint8_t thresh, int32_t dx,
                              int32_t dy, uint32_t filter_size, uint8_t bit_depth);
void oo_lf_narrow_filter(oxav1_plane_i32 *buf, int32_t x, int32_t y, int32_t dx, int32_t dy,
                         int hev_mask, uint8_t bit_depth);
void oo_lf_wide_filter(oxav1_plane_i32 *buf, int32_t x, int32_t y, uint8_t plane, int32_t dx,
                       int32_t dy, uint32_t log2_size);
int32_t oo_get_plane_residual_size(uint32_t subsize, uint8_t plane, uint8_t sx, uint8_t sy);
int32_t oo_get_tx_size(uint8_t plane, uint32_t tx_size, uint32_t mi_size, uint8_t sx, uint8_t sy);

/* ---- §7.15 CDEF (cdef.rs) ---- */
void oo_cdef_frame(const oxav1_cdef_ctx *ctx, const oxav1_cdef_grids *g,
                   const oxav1_plane_i32 *src, uint32_t nsrc, oxav1_plane_i32 *dst,
                   uint32_t ndst);
void oo_cdef_direction(const oxav1_cdef_ctx *ctx, const oxav1_plane_i32 *src, uint32_t nsrc,
                       uint32_t r, uint32_t c, int32_t *y_dir, int32_t *var);
void oo_cdef_filter_block(const oxav1_cdef_ctx *ctx, const oxav1_plane_i32 *src, uint32_t nsrc,
                          oxav1_plane_i32 *dst, uint32_t ndst, uint8_t plane, uint32_t r,
                          uint32_t c, int32_t pri_str, int32_t sec_str, int32_t damping,
                          int32_t dir);
int32_t oo_cdef_constrain(int32_t diff, int32_t threshold, int32_t damping);

/* ---- encoder CDEF election (encoder/cdef_elect.rs), oo_elect.c ---- */
int oo_cdef_elect(const oxav1_cdef_elect_input *in, oxav1_cdef_election *out);
void oo_cdef_run_units(const oxav1_cdef_elect_input *in, const oxav1_cdef_params *params, int which,
                       uint64_t *ssd);

/* ---- encoder loop-restoration election (encoder/lr_elect.rs), oo_lr_elect.c ---- */
int oo_lr_elect(const oxav1_lr_elect_input *in, oxav1_lr_election *out);
int oo_lr_apply_plan(const oxav1_lr_election *plan, const oxav1_lr_elect_input *in, uint16_t *const recon[3],
                     uint64_t *ssd);
/* The unit's output rects (x, y, w, h per rect, in push order); returns
 * the count (at most cap are written). */
uint32_t oo_lr_unit_rects(const oxav1_lr_elect_input *in, uint8_t plane, uint32_t ur, uint32_t uc,
                          uint32_t *rects, uint32_t cap);
/* fit_wiener over the rects; trace[s] = the fitted tap vector after step s
 * (round 0 horizontal, round 0 vertical, round 1 horizontal, round 1
 * vertical), taps = [vertical, horizontal] quantised. */
void oo_lr_fit_wiener(const int32_t *cdef, const uint16_t *src, uint32_t plane_w, uint32_t plane_h,
                      const uint32_t *rects, uint32_t nrects, int first_coeff, int32_t taps[2][3],
                      double trace[4][3]);
/* The counting symbol writer (encoder/symbol_writer.rs:168-345). */
uint32_t oo_log2_frac256(uint32_t x);
int64_t oo_recenter(int64_t r, int64_t v);
/* cost256 of a symbol sequence from range: kind 0 = symbol (value) over
 * cdf (n symbols), 1 = literal (n bits), 2 = subexp (num_syms = n, k),
 * 3 = signed subexp with reference (low .. high = n, k, r) — one element. */
uint64_t oo_cw_cost(uint32_t range, int kind, int64_t value, int64_t n, uint32_t k, int64_t low, int64_t r,
                    const uint16_t *cdf);
/* reps writes of one symbol over a fixed CDF, from range. */
uint64_t oo_cw_symbol_run(uint32_t range, uint32_t symbol, const uint16_t *cdf, uint32_t n, uint32_t reps);
/* The bools one signed subexp code writes (at most cap recorded), and its
 * cost256 from range 0x8000. */
uint32_t oo_cw_signed_ref_bits(int64_t low, int64_t high, uint32_t k, int64_t r, int64_t v, uint8_t *bits,
                               uint32_t cap, uint64_t *cost);
/* price_unit (lr_elect.rs:537-545) from the reference state ref_wiener
 * [2][3] / ref_xqd[2] of the plane: the SWITCHABLE symbol and the unit's
 * coefficients, cost256 from range 0x8000. */
uint64_t oo_lr_price_unit(const int32_t ref_wiener[2][3], const int32_t ref_xqd[2], int plane,
                          const oxav1_lr_unit *u, const uint16_t rt_cdf[4]);

/* ---- encoder superres election helpers (encoder/superres_elect.rs), oo_elect.c ---- */
uint32_t oo_superres_coded_width(uint32_t upscaled_width, uint32_t denom);
uint32_t oo_superres_candidate_denoms(uint32_t width, uint32_t *out);
int oo_superres_probe(const uint16_t *y, uint32_t w, uint32_t h, uint8_t bit_depth);
void oo_superres_downscale(const uint16_t *const src[3], uint32_t width, uint32_t height, uint8_t ssx, uint8_t ssy,
                           uint8_t num_planes, uint32_t coded_width, uint16_t *const dst[3]);
int oo_superres_upscale_recon(const uint16_t *const planes[3], uint32_t width, uint32_t height,
                              uint32_t upscaled_width, uint32_t mi_cols, uint8_t bit_depth, uint8_t ssx, uint8_t ssy,
                              uint8_t num_planes, uint16_t *const out[3]);

/* ---- encoder film-grain election helpers (encoder/film_grain_elect.rs), oo_elect.c ---- */
void oo_fg_denoise(const uint16_t *const src[3], uint32_t width, uint32_t height, uint8_t ssx, uint8_t ssy,
                   uint8_t num_planes, uint16_t *const dst[3]);
void oo_fg_noise_estimate(const uint16_t *const input[3], const uint16_t *const denoised[3], uint32_t width,
                          uint32_t height, uint8_t ssx, uint8_t ssy, uint8_t num_planes, uint8_t bit_depth,
                          oxav1_noise_estimate *out);
uint64_t oo_fg_plane_sigma16(const uint16_t *a, const uint16_t *b, size_t n, uint8_t bit_depth);
void oo_fg_apply_grain(const oxav1_film_grain_params *fg, const uint16_t *const planes[3], uint32_t width,
                       uint32_t height, uint8_t bit_depth, uint8_t ssx, uint8_t ssy, uint8_t num_planes,
                       uint8_t matrix_coefficients, uint16_t *const out[3]);

/* ---- §7.16 superres (superres.rs) ---- */
int oo_upscale_frame(const oxav1_sr_ctx *ctx, const oxav1_plane_i32 *in, uint32_t nin,
                     oxav1_plane_i32 *out, uint32_t nout);
int32_t oo_upscale_sample(const int32_t *row, uint32_t row_len, int32_t src_x_px,
                          uint32_t subpel, int32_t min_x, int32_t max_x, uint8_t bit_depth);

/* ---- §7.17 loop restoration (loop_restoration.rs) ---- */
typedef struct oo_lr_geom {
    uint32_t unit_row, unit_col, x, y, w, h;
    int32_t stripe_start_y, stripe_end_y, plane_end_x, plane_end_y;
} oo_lr_geom;
void oo_loop_restoration_frame(const oxav1_lr_ctx *ctx, const oxav1_lr_grids *g,
                               const oxav1_plane_i32 *curr, uint32_t ncurr,
                               const oxav1_plane_i32 *cdef, uint32_t ncdef,
                               oxav1_plane_i32 *lr, uint32_t nlr);
uint32_t oo_count_units_in_frame(uint32_t unit_size, uint32_t frame_size);
oo_lr_geom oo_lr_derive_block_geometry(const oxav1_lr_ctx *ctx, uint8_t plane, uint32_t row,
                                       uint32_t col);
int32_t oo_lr_get_source_sample(const oxav1_plane_i32 *curr, uint32_t ncurr,
                                const oxav1_plane_i32 *cdef, uint32_t ncdef, uint8_t plane,
                                int32_t x, int32_t y, const oo_lr_geom *geom);
void oo_wiener_coefficients(const int32_t coeff[3], int32_t filter[7]);
void oo_lr_restore_block(const oxav1_lr_ctx *ctx, const oxav1_lr_grids *g, const oxav1_plane_i32 *curr,
                         uint32_t ncurr, const oxav1_plane_i32 *cdef, uint32_t ncdef, oxav1_plane_i32 *lr,
                         uint32_t nlr, uint8_t plane, uint32_t row, uint32_t col);
/* box_filter (loop_restoration.rs:594-729); out has geom->h*geom->w entries. */
void oo_lr_box_filter(const oxav1_plane_i32 *curr, uint32_t ncurr, const oxav1_plane_i32 *cdef,
                      uint32_t ncdef, uint8_t plane, const oo_lr_geom *geom, uint8_t bit_depth,
                      int32_t r, int32_t eps, uint8_t pass, int32_t *out);

/* ---- §7.18.3 film grain (film_grain.rs) ---- */
uint32_t oo_fg_random(uint16_t *state, uint32_t bits);
/* luma: 73*82; cb/cr: chroma_h*chroma_w (<= 73*82). Returns chroma_w|chroma_h<<16. */
uint32_t oo_fg_generate_grain(const oxav1_film_grain_params *fg, uint8_t bit_depth, uint8_t sx,
                              uint8_t sy, int32_t *luma, int32_t *cb, int32_t *cr);
void oo_fg_scaling_lookup_init(const oxav1_film_grain_params *fg, uint8_t num_planes,
                               uint8_t lut[3][256]);
int32_t oo_fg_scale_lut(const uint8_t lut[3][256], uint8_t plane, int32_t index,
                        uint8_t bit_depth);
void oo_film_grain_synthesis(const oxav1_film_grain_params *fg, uint8_t bit_depth,
                             uint8_t num_planes, uint8_t sx, uint8_t sy,
                             uint8_t matrix_coefficients, oxav1_plane_i32 *planes, uint32_t n);

/* ---- Chain (frame_driver.rs:790-1116) ---- */
/*
 * Runs deblock -> CDEF -> superres -> LR -> crop -> (store) -> film grain on
 * the padded pre-filter planes `rec` (modified: they end as the deblocked
 * frame).  ref_out receives the cropped pre-grain planes, out the planes
 * after film grain (both UpscaledWidth x FrameHeight, subsampled).
 * stage_ns (optional, 5 entries) receives per-stage wall time.
 */
typedef struct oo_chain_grids {
    oxav1_lf_grids lf;
    oxav1_cdef_grids cdef;
    oxav1_lr_grids lr;
} oo_chain_grids;
int oo_chain(const oxav1_frame_desc *desc, const oo_chain_grids *g, oxav1_plane_i32 *rec,
             oxav1_plane_i32 *ref_out, oxav1_plane_i32 *out, double *stage_ns);

/* Table checksum for the table tests: sum of (i+1)*value over the table. */
int64_t oo_table_checksum(const char *name);

#ifdef __cplusplus
}
#endif
#endif
