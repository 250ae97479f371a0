/*
 * oxav1_gen.h — seeded synthetic workload generator for the filter chain.
 *
 * Produces what the CPU decode stages (entropy decode + reconstruction)
 * hand to the post-filter chain: padded pre-filter planes, the per-mi grids
 * the LoopFilter/CDEF/LR contexts sample, and legal per-frame parameters.
 * Deterministic: SplitMix64 streams derived from one 64-bit seed.
 *
 * Geometry follows frame_header.rs:1742-1746 (FrameWidth from UpscaledWidth
 * and SuperresDenom, MiCols = 2*((FrameWidth+7)>>3)).  The block grid is a
 * partition quadtree from the superblock size down to 4x4 with square
 * (and optionally HORZ/VERT) partitions and transform sizes tiled inside
 * each block, so every edge footprint is legal.
 */
#ifndef OXAV1_GEN_H
#define OXAV1_GEN_H

#include "oxav1_pf.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oxav1_gen_config {
    uint32_t upscaled_width;
    uint32_t frame_height;
    uint8_t bit_depth;       /* 8, 10, 12 */
    uint8_t subsampling_x;
    uint8_t subsampling_y;
    uint8_t mono;            /* 1 = one plane */
    uint8_t sb128;           /* superblock 128x128 (else 64x64) */
    uint8_t superres_denom;  /* 8 = no superres, 9..16 */
    uint8_t lr_unit_shift;   /* luma unit = 64 << shift (0..2) */
    uint8_t lr_uv_shift;     /* chroma unit = luma >> shift (only 4:2:0) */
    uint8_t enable_lf;
    uint8_t enable_cdef;
    uint8_t enable_lr;
    uint8_t enable_fg;
    uint8_t fg_overlap;
    uint8_t fg_num_y_points; /* 0..14 */
    uint8_t fg_ar_lag;       /* 0..3 */
    uint8_t fg_chroma_points;/* 0..10 per chroma plane (4:2:0 needs y points) */
    uint8_t noise_amp;       /* uniform noise +- amp around the gradient (bd scale) */
    uint8_t rect_partitions; /* allow HORZ/VERT partitions */
    uint8_t exercise_segments; /* random segment ids + ALT_LF features */
    uint8_t exercise_delta_lf; /* random per-mi delta_lf, delta_lf_multi drawn */
    uint8_t cdef_skip_units;   /* some 64x64 units get cdef_idx = -1 */
    uint8_t lr_types;          /* bit0 NONE, bit1 WIENER, bit2 SGRPROJ allowed */
    uint8_t decoder_geometry;  /* the block and transform geometry a decoder emits:
                                  HORZ_A/B, VERT_A/B and 4:1 (HORZ_4/VERT_4)
                                  partitions, transforms up to Max_Tx_Size_Rect
                                  (the 64-sample family), intra tx_depth splits
                                  and var-tx split trees in non-skip inter blocks */
} oxav1_gen_config;

/* One generated frame; all buffers are owned by the object. */
typedef struct oxav1_gen_frame {
    oxav1_frame_desc desc;
    /* padded pre-filter planes ((mi_rows*4)>>sy x (mi_cols*4)>>sx) */
    int32_t *planes[3];
    uint32_t plane_rows[3];
    uint32_t plane_cols[3];
    /* per-mi closure grids (mi_rows*mi_cols) */
    uint8_t *is_intra, *skip, *ref_frame, *mode, *segment_id, *mi_size;
    int8_t *delta_lf;          /* 4 per mi */
    uint8_t *lf_tx_size[3];    /* per plane grid */
    uint32_t tx_rows[3], tx_cols[3];
    uint32_t *lf_mi;           /* packed words (OXAV1_MI_* layout) */
    int8_t *cdef_idx;          /* per 64x64 unit */
    uint32_t cdef_rows, cdef_cols;
    oxav1_lr_unit *lr_units[3];
    uint32_t lr_unit_rows[3], lr_unit_cols[3];
} oxav1_gen_frame;

/* Defaults for one of the five benchmark configurations (0..4); for 3/4
 * the resolution / depth / subsampling draws are made from `seed`. */
int oxav1_gen_default_config(int cfg, uint64_t seed, oxav1_gen_config *out);
int oxav1_gen_frame_create(const oxav1_gen_config *cfg, uint64_t seed, oxav1_gen_frame **out);
void oxav1_gen_frame_destroy(oxav1_gen_frame *f);
/* SplitMix64 step, exported so callers derive per-frame seeds identically. */
uint64_t oxav1_gen_splitmix64(uint64_t x);

#ifdef __cplusplus
}
#endif
#endif
