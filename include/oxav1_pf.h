/*
 * oxav1_pf.h — C ABI of the MI355X (gfx950) AV1 post-reconstruction filter
 * chain: §7.14 deblocking, §7.15 CDEF, §7.16 superres, §7.17 loop
 * restoration and §7.18.3 film-grain synthesis.
 *
 * Two layers are exported:
 *
 *  1. Drop-in entry points (oxav1_pf_loop_filter_frame, _cdef_frame,
 *     _upscale_frame, _loop_restoration_frame, _film_grain_synthesis).
 *     Each replaces one public Rust function of oxideav-av1 with the same
 *     argument order and meaning.  Planes are the reference's PlaneBuffer
 *     (host int32 samples, stride = cols).  The per-block closures of the
 *     reference contexts become dense grids sampled on the caller's thread.
 *     The call stages the planes to HBM, runs the HIP kernels and returns
 *     after the results are back in the caller's buffers.
 *
 *  2. Device-resident engine (oxav1_pf_engine_*): whole-chain jobs on
 *     u8/u16 planes already in HBM, plain (one frame per launch) or batched
 *     (one persistent launch per stage over many frames, hipGraph replay).
 *
 * Every function returns an int status (OXAV1_OK or an error code below)
 * and never aborts.  No torch types appear in this interface.
 */
#ifndef OXAV1_PF_H
#define OXAV1_PF_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------- */
/* Status codes                                                            */
/* ---------------------------------------------------------------------- */
#define OXAV1_OK 0
/* The three SuperresError variants (superres.rs:217-230). */
#define OXAV1_ERR_UPSCALED_NOT_LARGER 1
#define OXAV1_ERR_PLANE_SHAPE_MISMATCH 2
#define OXAV1_ERR_PLANE_COUNT_MISMATCH 3
/* Conditions the reference cannot express (it would panic or clamp). */
#define OXAV1_ERR_INVALID_ARG 10
#define OXAV1_ERR_ALIAS 11
#define OXAV1_ERR_HIP 12
#define OXAV1_ERR_OUT_OF_MEMORY 13
#define OXAV1_ERR_ADMISSION_REFUSED 14
#define OXAV1_ERR_UNSUPPORTED 15
#define OXAV1_ERR_NO_DEVICE 16
#define OXAV1_ERR_BUSY 17
/* the frame's intermediates exceed the allocation-free entry point's
 * scratch (OXAV1_NOALLOC_SCRATCH_BYTES) */
#define OXAV1_ERR_SCRATCH_TOO_SMALL 18

/* ---------------------------------------------------------------------- */
/* Shared types                                                            */
/* ---------------------------------------------------------------------- */

/* PlaneBuffer (loop_filter.rs:119-126): row-major, stride = cols. */
typedef struct oxav1_plane_i32 {
    int32_t *samples;
    uint32_t rows;
    uint32_t cols;
} oxav1_plane_i32;

/* BLOCK_* ordinals (cdf.rs:14723-14770); 22 = BLOCK_INVALID. */
#define OXAV1_BLOCK_INVALID 22
/* FrameRestorationType (uncompressed_header_tail.rs:1019-1030). */
#define OXAV1_RESTORE_NONE 0
#define OXAV1_RESTORE_WIENER 1
#define OXAV1_RESTORE_SGRPROJ 2
#define OXAV1_RESTORE_SWITCHABLE 3

/* ---------------------------------------------------------------------- */
/* §7.14 deblocking                                                        */
/* ---------------------------------------------------------------------- */

/* Scalar fields of LoopFilterFrameContext (loop_filter.rs:167-201). */
typedef struct oxav1_lf_ctx {
    uint8_t loop_filter_level[4];
    uint8_t loop_filter_sharpness;
    uint8_t loop_filter_delta_enabled;
    uint8_t delta_lf_multi;
    uint8_t num_planes;
    int8_t loop_filter_ref_deltas[8];
    int8_t loop_filter_mode_deltas[2];
    uint8_t bit_depth;
    uint8_t subsampling_x;
    uint8_t subsampling_y;
    uint8_t reserved0;
    uint32_t mi_rows;
    uint32_t mi_cols;
    uint32_t frame_width;  /* coded FrameWidth (frame_driver.rs:853) */
    uint32_t frame_height;
} oxav1_lf_ctx;

/*
 * The closures of LoopFilterFrameContext (loop_filter.rs:202-237) sampled
 * into dense grids.  Per-mi grids are mi_rows*mi_cols, row-major.
 *   is_intra(r,c), skip(r,c), ref_frame(r,c), mode(r,c), segment_id(r,c),
 *   mi_size(r,c): one byte each.
 *   delta_lf(r,c,idx): 4 signed bytes per mi cell; NULL means all zero.
 *   segment_id: NULL means 255 (no segment) everywhere.
 *   lf_tx_size(plane, r, c): plane p grid is ((mi_rows+sy)>>sy) rows by
 *   ((mi_cols+sx)>>sx) cols, sx/sy = 0 for plane 0.
 *   seg_feature_active/data(segment, feature): 8 segments x 8 features;
 *   segment >= 8 reads inactive/0 (the bridge at cdf.rs:20514-20530).
 */
typedef struct oxav1_lf_grids {
    const uint8_t *is_intra;
    const uint8_t *skip;
    const uint8_t *ref_frame;
    const uint8_t *mode;
    const uint8_t *segment_id;
    const int8_t *delta_lf;
    const uint8_t *mi_size;
    const uint8_t *lf_tx_size[3];
    uint8_t seg_feature_active[8][8];
    int16_t seg_feature_data[8][8];
} oxav1_lf_grids;

/* Replaces loop_filter_frame (loop_filter.rs:271). In place on planes. */
int oxav1_pf_loop_filter_frame(const oxav1_lf_ctx *ctx, const oxav1_lf_grids *grids,
                               oxav1_plane_i32 *planes, uint32_t num_plane_bufs);

/* ---------------------------------------------------------------------- */
/* §7.15 CDEF                                                              */
/* ---------------------------------------------------------------------- */

/* CdefParams (uncompressed_header_tail.rs:876-900). */
typedef struct oxav1_cdef_params {
    uint8_t cdef_damping;
    uint8_t cdef_bits;
    uint8_t y_pri_strength[8];
    uint8_t y_sec_strength[8];
    uint8_t uv_pri_strength[8];
    uint8_t uv_sec_strength[8];
} oxav1_cdef_params;

/* Scalar fields of CdefFrameContext (cdef.rs:114-135). */
typedef struct oxav1_cdef_ctx {
    uint32_t mi_rows;
    uint32_t mi_cols;
    uint8_t num_planes;
    uint8_t bit_depth;
    uint8_t subsampling_x;
    uint8_t subsampling_y;
    oxav1_cdef_params params;
} oxav1_cdef_ctx;

/*
 * cdef_idx(r,c) is only queried at 64x64 anchors (cdef.rs:193-195), so it is
 * one signed byte per 64x64 unit: ((mi_rows+15)>>4) x ((mi_cols+15)>>4).
 * skip(r,c) is mi_rows*mi_cols bytes.
 */
typedef struct oxav1_cdef_grids {
    const int8_t *cdef_idx;
    const uint8_t *skip;
} oxav1_cdef_grids;

/* Replaces cdef_frame (cdef.rs:180). src and dst must not alias. */
int oxav1_pf_cdef_frame(const oxav1_cdef_ctx *ctx, const oxav1_cdef_grids *grids,
                        const oxav1_plane_i32 *src, uint32_t num_src,
                        oxav1_plane_i32 *dst, uint32_t num_dst);

/* ---------------------------------------------------------------------- */
/* §7.16 superres                                                          */
/* ---------------------------------------------------------------------- */

/* SuperresFrameContext (superres.rs:239-264). */
typedef struct oxav1_sr_ctx {
    uint8_t use_superres;
    uint8_t num_planes;
    uint8_t bit_depth;
    uint8_t subsampling_x;
    uint8_t subsampling_y;
    uint8_t reserved[3];
    uint32_t frame_width;
    uint32_t upscaled_width;
    uint32_t frame_height;
    uint32_t mi_cols;
} oxav1_sr_ctx;

/* Replaces upscale_frame (superres.rs:282); returns the SuperresError codes.
 * Like the reference, planes before a failing plane are already written. */
int oxav1_pf_upscale_frame(const oxav1_sr_ctx *ctx, const oxav1_plane_i32 *in, uint32_t num_in,
                           oxav1_plane_i32 *out, uint32_t num_out);

/* ---------------------------------------------------------------------- */
/* §7.17 loop restoration                                                  */
/* ---------------------------------------------------------------------- */

/* One restoration unit: LrType / LrWiener / LrSgrSet / LrSgrXqd
 * (loop_restoration.rs:217-233). wiener[pass][i], pass 0 = vertical. */
typedef struct oxav1_lr_unit {
    uint8_t type;    /* OXAV1_RESTORE_* */
    uint8_t sgr_set; /* 0..15 */
    int16_t wiener[2][3];
    int16_t sgr_xqd[2];
} oxav1_lr_unit;

/* LoopRestorationFrameContext scalars + LrParams (loop_restoration.rs:194-216,
 * uncompressed_header_tail.rs:1079-1109). */
typedef struct oxav1_lr_ctx {
    uint32_t mi_rows;
    uint32_t mi_cols;
    uint8_t num_planes;
    uint8_t bit_depth;
    uint8_t subsampling_x;
    uint8_t subsampling_y;
    uint32_t frame_height;
    uint32_t upscaled_width;
    uint8_t frame_restoration_type[3];
    uint8_t uses_lr;
    uint32_t loop_restoration_size[3];
} oxav1_lr_ctx;

/* Per-plane unit grids, unit_rows x unit_cols with
 * unit_rows = count_units_in_frame(size, Round2(FrameHeight, sy)) and
 * unit_cols = count_units_in_frame(size, Round2(UpscaledWidth, sx))
 * (loop_restoration.rs:395-398). NULL = every unit RESTORE_NONE. */
typedef struct oxav1_lr_grids {
    const oxav1_lr_unit *units[3];
} oxav1_lr_grids;

/* Replaces loop_restoration_frame (loop_restoration.rs:268).
 * curr = pre-CDEF (upscaled) planes, cdef = CDEF output, lr = output. */
int oxav1_pf_loop_restoration_frame(const oxav1_lr_ctx *ctx, const oxav1_lr_grids *grids,
                                    const oxav1_plane_i32 *curr, uint32_t num_curr,
                                    const oxav1_plane_i32 *cdef, uint32_t num_cdef,
                                    oxav1_plane_i32 *lr, uint32_t num_lr);

/* count_units_in_frame (loop_restoration.rs:910-915). */
uint32_t oxav1_pf_count_units_in_frame(uint32_t unit_size, uint32_t frame_size);

/* ---------------------------------------------------------------------- */
/* §7.18.3 film grain                                                      */
/* ---------------------------------------------------------------------- */

/* FilmGrainParams (uncompressed_header_tail.rs:1763-1834). grain_scaling is
 * already +8 and ar_coeff_shift already +6, as in the reference. */
typedef struct oxav1_film_grain_params {
    uint8_t apply_grain;
    uint8_t update_grain;
    uint16_t grain_seed;
    uint8_t num_y_points;
    uint8_t point_y_value[14];
    uint8_t point_y_scaling[14];
    uint8_t chroma_scaling_from_luma;
    uint8_t num_cb_points;
    uint8_t point_cb_value[10];
    uint8_t point_cb_scaling[10];
    uint8_t num_cr_points;
    uint8_t point_cr_value[10];
    uint8_t point_cr_scaling[10];
    uint8_t grain_scaling;
    uint8_t ar_coeff_lag;
    uint8_t ar_coeffs_y_plus_128[24];
    uint8_t ar_coeffs_cb_plus_128[25];
    uint8_t ar_coeffs_cr_plus_128[25];
    uint8_t ar_coeff_shift;
    uint8_t grain_scale_shift;
    uint8_t cb_mult;
    uint8_t cb_luma_mult;
    uint16_t cb_offset;
    uint8_t cr_mult;
    uint8_t cr_luma_mult;
    uint16_t cr_offset;
    uint8_t overlap_flag;
    uint8_t clip_to_restricted_range;
} oxav1_film_grain_params;

/* Replaces film_grain_synthesis (film_grain.rs:848). In place on planes
 * (the cropped output planes; planes[0] gives w/h as in film_grain.rs:585). */
int oxav1_pf_film_grain_synthesis(const oxav1_film_grain_params *fg, uint8_t bit_depth,
                                  uint8_t num_planes, uint8_t subsampling_x,
                                  uint8_t subsampling_y, uint8_t matrix_coefficients,
                                  oxav1_plane_i32 *planes, uint32_t num_plane_bufs);

/* Staging of the five calls above: every sample the decoder passes is a
 * Clip1-ed value in [0, 2^bit_depth), so a call whose int32 input planes
 * all lie in that range stages them as u8 (8-bit) or u16 planes, converted
 * on the host while copying, and runs the u8/u16 kernels; the results are
 * widened back to int32.  A call with any sample out of range stages int32
 * planes and runs the int32 kernels.  Results are the same either way.
 * enable = 0 forces the int32 path for every call (default: 1).  No
 * counterpart in the reference (its planes are always i32). */
int oxav1_pf_dropin_narrowing(int enable);

/* ---------------------------------------------------------------------- */
/* Library / device                                                        */
/* ---------------------------------------------------------------------- */

/* Initialise the HIP runtime on the current device. Optional: every entry
 * point calls it lazily. Returns OXAV1_ERR_NO_DEVICE without a GPU. */
int oxav1_pf_init(void);
/* Human-readable text for a status code (static storage). */
const char *oxav1_pf_strerror(int status);
/* Text of the last HIP error seen by this thread (static per-thread). */
const char *oxav1_pf_last_hip_error(void);
/* Sum of (i+1)*value over one of the library's constant tables (names as in
 * the reference: "GAUSSIAN_SEQUENCE", "UPSCALE_FILTER", "SGR_PARAMS", ...);
 * INT64_MIN for an unknown name. */
int64_t oxav1_pf_table_checksum(const char *name);
/* sizeof of an ABI struct by its type name ("oxav1_lf_ctx", ...), 0 if
 * unknown; lets foreign bindings verify their layouts. */
size_t oxav1_pf_struct_size(const char *name);

/* Device memory on the library's own HIP runtime, for callers that have no
 * allocator of their own (bindings, tests, benchmarks).  Synchronous. */
int oxav1_pf_dev_alloc(void **ptr, size_t bytes);
int oxav1_pf_dev_free(void *ptr);
/* Page-locked host memory (hipHostMalloc), for the host buffers of
 * oxav1_pf_server_submit_host: copies from and to it run asynchronously. */
int oxav1_pf_host_alloc(void **ptr, size_t bytes);
int oxav1_pf_host_free(void *ptr);
int oxav1_pf_dev_upload(void *dst, const void *src, size_t bytes);
int oxav1_pf_dev_download(void *dst, const void *src, size_t bytes);
int oxav1_pf_dev_fill(void *dst, int byte_value, size_t bytes);
/* Device properties: compute units and HBM bytes free / total. */
int oxav1_pf_dev_info(int *num_cus, size_t *free_bytes, size_t *total_bytes);
/* Device-to-device copy by a plain 16-byte-per-lane kernel (bytes % 16 == 0);
 * used to calibrate HBM traffic counters against a known byte count. */
int oxav1_pf_dev_copy_kernel(void *dst, const void *src, size_t bytes, int repeats);

/* ---------------------------------------------------------------------- */
/* Device-resident chain engine                                            */
/* ---------------------------------------------------------------------- */

/*
 * Packed per-mi side information for the deblocking kernels, one uint32 per
 * mi cell (the LoopFilterFrameContext closures in a single word):
 *   bits 0-4 mi_size, bit 5 skip, bit 6 is_intra, bits 8-11 ref_frame,
 *   bits 12-16 mode, bits 24-31 segment_id.
 */
#define OXAV1_MI_SIZE_SHIFT 0
#define OXAV1_MI_SKIP_BIT 5
#define OXAV1_MI_INTRA_BIT 6
#define OXAV1_MI_REF_SHIFT 8
#define OXAV1_MI_MODE_SHIFT 12
#define OXAV1_MI_SEG_SHIFT 24

/* Frame-level parameters of one chain job (the gates of
 * frame_driver.rs:832-1096 plus the per-stage parameter structs). */
typedef struct oxav1_frame_desc {
    uint32_t frame_width;    /* coded (post-superres-downscale) width */
    uint32_t frame_height;
    uint32_t upscaled_width; /* == frame_width without superres */
    uint32_t mi_rows;
    uint32_t mi_cols;
    uint8_t bit_depth;
    uint8_t subsampling_x;
    uint8_t subsampling_y;
    uint8_t num_planes;
    uint8_t matrix_coefficients;
    uint8_t use_superres;
    uint8_t apply_lf;   /* (lvl[0]||lvl[1]) && !coded_lossless && !allow_intrabc */
    uint8_t apply_cdef; /* !coded_lossless && !allow_intrabc && enable_cdef */
    oxav1_lf_ctx lf;
    oxav1_cdef_params cdef;
    oxav1_lr_ctx lr;
    oxav1_film_grain_params fg;
    uint8_t seg_feature_active[8][8];
    int16_t seg_feature_data[8][8];
} oxav1_frame_desc;

/* Device pointers of one job. Planes are u8 (8-bit) or u16 (10/12-bit);
 * pitches are in samples.  src is the padded pre-filter CurrFrame
 * ((mi_rows*4)>>sy x (mi_cols*4)>>sx), ref_out the cropped pre-grain frame
 * (UpscaledWidth x FrameHeight, subsampled), out the film-grain output
 * (may equal ref_out's buffers only when film grain is off). */
typedef struct oxav1_frame_dev {
    const void *src[3];
    uint32_t src_pitch[3];
    void *ref_out[3];
    uint32_t ref_pitch[3];
    void *out[3];
    uint32_t out_pitch[3];
    const uint32_t *lf_mi;       /* mi_rows*mi_cols packed words */
    const int8_t *delta_lf;      /* mi_rows*mi_cols*4 or NULL */
    const uint8_t *lf_tx_size[3];
    const int8_t *cdef_idx;      /* per 64x64 unit */
    const oxav1_lr_unit *lr_units[3];
} oxav1_frame_dev;

/* The block walker's per-mi grids (device pointers, one byte per mi cell,
 * row-major MiRows x MiCols), as cdf.rs:20320-20678 reads them for the
 * loop-filter closures: is_inter, skip, RefFrame[0], YMode, segment_id,
 * MiSize, TxSizes and InterTxSizes. */
typedef struct oxav1_walker_grids {
    const uint8_t *is_inter;
    const uint8_t *skip;
    const uint8_t *ref_frame;
    const uint8_t *mode;
    const uint8_t *segment_id;
    const uint8_t *mi_size;
    const uint8_t *tx_size;
    const uint8_t *inter_tx_size;
} oxav1_walker_grids;

/* Build the engine's side information on the device from the walker's
 * grids, asynchronously on hip_stream (NULL = the default stream): lf_mi
 * (MiRows*MiCols packed words, OXAV1_MI_* layout) and lf_tx_size[p]
 * (loopfilter_tx_size_at, cdf.rs:20559-20593: luma from InterTxSizes for
 * inter cells and TxSizes for intra cells; chroma by get_tx_size of the
 * co-located luma mi, cdf.rs:12976-13012, at ((MiRows+sy)>>sy) x
 * ((MiCols+sx)>>sx)).  Replaces the host-side closure sampling. */
int oxav1_pf_build_grids(const oxav1_walker_grids *g, uint32_t mi_rows, uint32_t mi_cols, uint8_t num_planes,
                         uint8_t subsampling_x, uint8_t subsampling_y, uint32_t *lf_mi,
                         uint8_t *const lf_tx_size[3], void *hip_stream);

typedef struct oxav1_pf_engine oxav1_pf_engine;

/* Create an engine bound to the current device and a HIP stream (NULL =
 * the engine creates its own default-priority stream). max_frames bounds
 * a batch; intermediate buffers are allocated per frame slot by
 * oxav1_pf_engine_prepare and grow on demand. */
int oxav1_pf_engine_create(oxav1_pf_engine **out, void *hip_stream, uint32_t max_frames);
int oxav1_pf_engine_destroy(oxav1_pf_engine *e);
/* Admit a batch of n frames: validate, size the per-frame intermediates
 * (deblocked, CDEF, upscaled planes, film-grain scratch) and upload the
 * device job table.  Synchronous.  Returns OXAV1_ERR_ADMISSION_REFUSED when
 * n exceeds max_frames and OXAV1_ERR_OUT_OF_MEMORY when HBM is exhausted. */
int oxav1_pf_engine_prepare(oxav1_pf_engine *e, const oxav1_frame_desc *desc,
                            const oxav1_frame_dev *dev, uint32_t n);
/* Run the prepared batch, asynchronously on the engine's stream.
 * mode 0: plain, one launch per stage per frame;
 * mode 1: batched, one persistent launch per stage pulling
 *         (frame, plane, tile) items from a device counter;
 * mode 2: mode 1 captured once in a hipGraph and replayed. */
int oxav1_pf_engine_run(oxav1_pf_engine *e, int mode);
/* prepare(n = 1) + run(mode 0). */
int oxav1_pf_engine_run_frame(oxav1_pf_engine *e, const oxav1_frame_desc *desc,
                              const oxav1_frame_dev *dev);
/* prepare(n) + run(use_graph ? 2 : 1). */
int oxav1_pf_engine_run_batch(oxav1_pf_engine *e, const oxav1_frame_desc *desc,
                              const oxav1_frame_dev *dev, uint32_t n, int use_graph);
/* The engine's HIP stream (hipStream_t). */
void *oxav1_pf_engine_stream(oxav1_pf_engine *e);
/* Wait for all work on the engine's stream. */
int oxav1_pf_engine_sync(oxav1_pf_engine *e);
/* Launch shape knobs, chosen per filter and per sample size by the
 * autotuner (oxav1/autotune.py) and applied from the next prepare.  Index
 * [0] applies to 8-bit frames (batches), [1] to 10/12-bit.  Zero fields
 * keep the defaults.  None of them changes any output byte. */
typedef struct oxav1_tuning {
    uint16_t blocks_per_cu[5][2]; /* persistent workgroups per CU for lf, cdef, sr, lr, fg
                                     (0 = as many as fit); a batch uses the row of the
                                     sample size holding most of its pixels */
    uint8_t waves_per_eu[5][2];   /* register budget of each stage's kernel: the compiled
                                     variant that allows at least 6, 7 or 8 waves per SIMD
                                     (0 = the compiler's choice) */
    uint8_t sr_rows_per_item[2];  /* superres rows per work item (0 = 8, at most 64) */
    uint8_t lr_tile_w[2];         /* loop restoration tile width: 32, or 64 where it divides
                                     the unit (0 = widest that divides) */
    uint8_t sub_batches[2];       /* batched path split into this many sub-batches of balanced
                                     pixels on streams of their own, so stages of different
                                     frames overlap (0 or 1 = one batch, at most 4) */
    uint8_t items_per_claim[5][2]; /* work items a persistent workgroup takes per atomic on
                                     the stage's counter: 0 or 1 = one, up to 16, or 255 =
                                     no atomics (items blockIdx.x + k * gridDim.x) */
} oxav1_tuning;
int oxav1_pf_engine_set_tuning(oxav1_pf_engine *e, const oxav1_tuning *t);
/* Current tuning, the CU count and how many workgroups of each stage fit
 * on one CU, per waves_per_eu variant (0, 6, 7, 8); any pointer may be NULL. */
int oxav1_pf_engine_get_tuning(oxav1_pf_engine *e, oxav1_tuning *t, uint32_t *num_cus,
                               uint32_t fit_blocks_per_cu[5][4]);
/* OXAV1_OK when everything enqueued on the engine's stream has finished,
 * OXAV1_ERR_BUSY while work is in flight; never blocks. */
int oxav1_pf_engine_query(oxav1_pf_engine *e);
/* Allocation-free variant of oxav1_pf_engine_run_frame, for callers that
 * cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  Runs the plain chain (one launch per stage) over one frame on
 * hip_stream (NULL = the default stream), asynchronously, with the same
 * kernels and outputs as oxav1_pf_engine_run_frame.  The frame's
 * intermediates (deblocked, CDEF and upscaled planes, film-grain scratch)
 * and its job record live in a fixed device array of
 * OXAV1_NOALLOC_SCRATCH_BYTES per GPU, held from the library's device
 * set-up until the process ends; a frame whose intermediates do not fit
 * returns OXAV1_ERR_SCRATCH_TOO_SMALL before anything is enqueued (a
 * 3840x2160 10-bit 4:2:0 frame coded 2560 wide with superres and every
 * stage on needs 79.3 MiB).  Nothing is allocated and nothing blocks: the
 * call only enqueues kernels (the scratch is written by kernels, not by
 * memset or copy nodes).  All calls on one GPU share its array and leave
 * the last call's contents in it: calls that overlap on one GPU (on
 * different streams, or from different threads) race and must be
 * serialised by the caller. */
#define OXAV1_NOALLOC_SCRATCH_BYTES (80u << 20)
int oxav1_pf_run_frame_noalloc(const oxav1_frame_desc *desc, const oxav1_frame_dev *dev, void *hip_stream);
/* Size n frame slots for frames of desc's shape with every stage on, so
 * that prepare with frames no larger never allocates; *bytes (optional)
 * receives the engine's slot bytes. */
int oxav1_pf_engine_reserve(oxav1_pf_engine *e, const oxav1_frame_desc *desc, uint32_t n,
                            uint64_t *bytes);
/* Time the dominant kernel of the last runs: average duration in ms of the
 * named stage ("lf","cdef","sr","lr","fg") measured with HIP events on the
 * engine's stream when timing is enabled. */
int oxav1_pf_engine_set_timing(oxav1_pf_engine *e, int enable);
int oxav1_pf_engine_stage_ms(oxav1_pf_engine *e, const char *stage, double *avg_ms,
                             uint32_t *count);

/* ---------------------------------------------------------------------- */
/* Serving loop: concurrent decode sessions, continuous batching           */
/* ---------------------------------------------------------------------- */

/*
 * A server owns `lanes` engines, each on its own stream, and one HBM arena
 * from which sessions are admitted by footprint.  A session is a decode
 * stream of frames of one shape; it owns `ref_slots` post-filter reference
 * slots (the pre-grain frame store of frame_driver.rs:1044-1052, plus a
 * film-grain output per slot when its frames carry grain) and an optional
 * input region for its pre-filter frames, all resident in the arena until
 * the session closes.  Frames are submitted as their pre-filter data becomes
 * ready and wait in one FIFO; every poll retires finished lanes and hands the
 * frames ready at that moment (up to max_batch) to a free lane as one
 * batched launch per stage, so frames that arrive while others are on the
 * device join the next free lane and lanes overlap on the device.
 * One host thread drives a server (the calls are not thread-safe).
 */
typedef struct oxav1_pf_server oxav1_pf_server;

typedef struct oxav1_server_config {
    uint32_t lanes;              /* engines on their own streams (0 = 4) */
    uint32_t max_batch;          /* frames per lane launch (0 = 16) */
    uint64_t arena_bytes;        /* HBM for sessions (0 = half the free HBM) */
    const oxav1_frame_desc *largest; /* lane intermediates sized for it (NULL = grow on demand) */
} oxav1_server_config;

typedef struct oxav1_server_done {
    uint64_t ticket;
    uint32_t session;
    uint32_t slot;
    int32_t status;              /* OXAV1_OK or the failing call's status */
    uint32_t batch;              /* frames in the launch that carried it */
    double arrival_us;           /* server clock (oxav1_pf_server_now_us) */
    double launch_us;
    double done_us;              /* when a poll saw the lane finish */
} oxav1_server_done;

typedef struct oxav1_server_stats {
    uint32_t sessions_open;
    uint32_t lanes_busy;
    uint32_t frames_queued;
    uint32_t max_batch_seen;
    uint64_t frames_submitted;
    uint64_t frames_done;
    uint64_t launches;
    uint64_t arena_bytes;
    uint64_t arena_used;
    uint64_t lane_bytes;         /* lane intermediates reserved at create */
} oxav1_server_stats;

double oxav1_pf_server_now_us(void);
int oxav1_pf_server_create(oxav1_pf_server **out, const oxav1_server_config *cfg);
/* Waits for every lane, then frees the engines and the arena. */
int oxav1_pf_server_destroy(oxav1_pf_server *s);
/* Admit a session whose frames have `shape` (sizes, depth, subsampling,
 * superres and film-grain flags): reserves ref_slots reference slots plus
 * input_bytes of input region in the arena.  OXAV1_ERR_ADMISSION_REFUSED
 * when the footprint does not fit what the open sessions leave free;
 * *input (optional) receives the input region's device pointer. */
int oxav1_pf_server_open_session(oxav1_pf_server *s, const oxav1_frame_desc *shape, uint32_t ref_slots,
                                 uint64_t input_bytes, uint32_t *session, void **input,
                                 uint64_t *footprint);
/* OXAV1_ERR_BUSY while a frame of the session is queued or in flight. */
int oxav1_pf_server_close_session(oxav1_pf_server *s, uint32_t session);
/* Device pointer and pitch (samples) of a slot's plane: which = 0 for the
 * pre-grain reference frame, 1 for the film-grain output. */
int oxav1_pf_server_slot(oxav1_pf_server *s, uint32_t session, uint32_t slot, uint32_t plane, int which,
                         void **ptr, uint32_t *pitch);
/* Queue one frame of a session; its outputs go to the session's slot
 * (dev's ref_out/out are replaced by the slot's).  The frame's shape must
 * match the session's.  OXAV1_ERR_BUSY when the slot still has a frame
 * queued or in flight. */
int oxav1_pf_server_submit(oxav1_pf_server *s, uint32_t session, const oxav1_frame_desc *desc,
                           const oxav1_frame_dev *dev, uint32_t slot, double arrival_us, uint64_t *ticket);
/* Queue the show_existing_frame output of a slot (frame_driver.rs:1489-1537):
 * film grain with `fg` (the slot's saved grain parameters) applied to the
 * slot's stored pre-grain frame, written to the slot's grain buffer; the
 * stored frame is not modified.  The session must have been opened with a
 * grain shape (fg.apply_grain set) so that its slots carry grain buffers.
 * Without grain (fg->apply_grain == 0) the stored frame is copied. */
int oxav1_pf_server_submit_existing(oxav1_pf_server *s, uint32_t session, uint32_t slot,
                                    const oxav1_film_grain_params *fg, double arrival_us, uint64_t *ticket);
/* One frame as the CPU stages of the decoder leave it, in host memory
 * (page-locked, oxav1_pf_host_alloc, so the copies are asynchronous): the
 * pre-filter planes (u8 at 8-bit, u16 above; the padded CurrFrame,
 * (mi_rows*4)>>sy x (mi_cols*4)>>sx, pitch in samples, 0 = tight) and the
 * side information laid out as in oxav1_frame_dev.  ref_out / out are
 * optional host destinations (cropped output shape, tight rows) of the
 * pre-grain frame and the film-grain output, filled before the frame is
 * reported done (frame_driver.rs:1098-1116 hands the narrowed output back to
 * the caller); NULL leaves the result only in the session's slot. */
typedef struct oxav1_frame_host {
    const void *src[3];
    uint32_t src_pitch[3];
    const uint32_t *lf_mi;
    const int8_t *delta_lf;      /* NULL = no delta_lf */
    const uint8_t *lf_tx_size[3];
    const int8_t *cdef_idx;
    const oxav1_lr_unit *lr_units[3];
    void *ref_out[3];
    void *out[3];
} oxav1_frame_host;

/* Queue one frame whose inputs are in host memory: the server copies them
 * into one of the session's input buffers (its input region, opened with
 * input_bytes, holds input_bytes / oxav1_pf_server_frame_input_bytes such
 * buffers) on its upload stream at once, so the copy overlaps the frames
 * already on the device; the lane that takes the frame waits for that copy,
 * and copies the results back to the host destinations after the chain.
 * OXAV1_ERR_BUSY when the slot is busy or every input buffer holds a frame
 * still queued or in flight. */
int oxav1_pf_server_submit_host(oxav1_pf_server *s, uint32_t session, const oxav1_frame_desc *desc,
                                const oxav1_frame_host *in, uint32_t slot, double arrival_us, uint64_t *ticket);
/* Device bytes one frame of this shape needs in a session's input region
 * for oxav1_pf_server_submit_host (planes and side information, 256-byte
 * aligned pieces). */
uint64_t oxav1_pf_server_frame_input_bytes(const oxav1_frame_desc *shape);

/* Retire finished lanes, launch ready frames on free lanes, and return up
 * to cap finished frames (the rest are kept for the next poll). */
int oxav1_pf_server_poll(oxav1_pf_server *s, oxav1_server_done *done, uint32_t cap, uint32_t *n_done);
int oxav1_pf_server_stats(oxav1_pf_server *s, oxav1_server_stats *st);

/* ---------------------------------------------------------------------- */
/* Encoder elections on the filter kernels                                 */
/* ---------------------------------------------------------------------- */

/* One CDEF strength candidate for one plane set (plane_set 0: luma,
 * 1: both chroma planes), as the encoder's search evaluates it
 * (encoder/cdef_elect.rs:180-190 `params_for`): damping and the set's
 * (pri, sec) strengths in strength slot 0, every other slot zero. */
typedef struct oxav1_cdef_cand {
    uint8_t damping;
    uint8_t plane_set;
    uint8_t pri;
    uint8_t sec;
} oxav1_cdef_cand;

/* One frame of a per-unit SSD job, all pointers in device memory. Planes
 * are u8 when sample_bytes == 1, u16 when 2; pitches in samples. rows/cols
 * are the plane buffers' shapes (the reference's PlaneBuffer rows/cols);
 * the 8x8 walk over the mi grid must cover them. cdef_idx has one byte per
 * 64x64 unit ((mi_rows+15)>>4 x (mi_cols+15)>>4), skip one per mi cell.
 * ssd receives ncand x n_units uint64 (candidate-major). */
typedef struct oxav1_cdef_ssd_frame {
    uint32_t mi_rows, mi_cols;
    const int8_t *cdef_idx;
    const uint8_t *skip;
    const void *recon[3];
    const void *input[3];
    uint32_t recon_pitch[3], input_pitch[3];
    uint32_t rows[3], cols[3];
    uint8_t bit_depth, subsampling_x, subsampling_y, num_planes;
    uint8_t sample_bytes, pad[3];
    uint64_t *ssd;
} oxav1_cdef_ssd_frame;

/* Replaces the per-candidate `run_units` pass of the CDEF election
 * (encoder/cdef_elect.rs:192-240): for every candidate and every 64x64
 * unit, the SSD against `input` of the unit's samples of the candidate's
 * plane set after one §7.15 run over `recon` (units whose cdef_idx is 0
 * filter with the candidate; other ids and skip blocks copy). The §7.15.2
 * direction search runs once per 8x8 block for all candidates. One launch
 * over all frames, asynchronous on `hip_stream` (NULL = the library
 * stream, synchronised before return). ncand <= 128. */
int oxav1_pf_cdef_unit_ssd(const oxav1_cdef_ssd_frame *frames, uint32_t nframes, const oxav1_cdef_cand *cands,
                           uint32_t ncand, void *hip_stream);

/* CdefElectInput (encoder/cdef_elect.rs:91-113). The walker's committed
 * grids are cdef_idx (one byte per 64x64 unit, -1 = no idx coded) and skip
 * (one byte per mi cell); planes are host u16 arrays of width x height
 * (luma) and chroma_w x chroma_h. */
typedef struct oxav1_cdef_elect_input {
    uint32_t mi_rows, mi_cols;
    const int8_t *cdef_idx;
    const uint8_t *skip;
    const uint16_t *input[3];
    const uint16_t *recon[3];
    uint32_t width, height, chroma_w, chroma_h;
    uint8_t bit_depth, subsampling_x, subsampling_y, num_planes;
    uint8_t max_bits, pad[3];
    uint64_t lambda;
} oxav1_cdef_elect_input;

/* CdefElection (encoder/cdef_elect.rs:72-89). elected = 0 is the
 * reference's None (no coded unit, or the unfiltered frame won). The
 * unit-id arrays are caller-owned, n_units = ((mi_rows+15)>>4) x
 * ((mi_cols+15)>>4) entries each, written when the plan exists. */
typedef struct oxav1_cdef_election {
    int32_t elected;
    int32_t has_frame_level;
    oxav1_cdef_params best;
    oxav1_cdef_params frame_level;
    uint64_t best_d, frame_level_d, base_d;
    int8_t *best_unit_idx;
    int8_t *frame_level_unit_idx;
} oxav1_cdef_election;

/* Replaces elect_cdef (encoder/cdef_elect.rs:136): the same candidate
 * sweeps, damping refinement, greedy per-unit set growth and D*256 +
 * lambda*R256 choice, with every candidate's per-unit SSD table computed
 * on the GPU (three launches of oxav1_pf_cdef_unit_ssd). Synchronous. */
int oxav1_pf_cdef_elect(const oxav1_cdef_elect_input *in, oxav1_cdef_election *out);

/* Superres election helpers (encoder/superres_elect.rs). Planes are host
 * u16 arrays, row-major without padding. */

/* superres_coded_width (superres_elect.rs:23-29): the §5.9.8 FrameWidth. */
uint32_t oxav1_pf_superres_coded_width(uint32_t upscaled_width, uint32_t denom);
/* candidate_denoms (superres_elect.rs:31-45): writes up to 8 denominators
 * to `denoms` and returns their number. */
uint32_t oxav1_pf_superres_candidate_denoms(uint32_t width, uint32_t *denoms);
/* superres_probe (superres_elect.rs:62-91): *probe = 1 when the mean
 * absolute horizontal luma second difference is under 10/16 of an 8-bit
 * step. Computed on the GPU. */
int oxav1_pf_superres_probe(const uint16_t *y, uint32_t width, uint32_t height, uint8_t bit_depth,
                            int32_t *probe);
/* downscale_width (superres_elect.rs:93-150): every plane low-passed
 * [1,2,1]/4 and resampled centre-aligned at 1/64-sample precision, luma
 * from width to coded_width columns, chroma from width >> ssx to
 * coded_width >> ssx (heights height and height >> ssy). coded_width must
 * be below width and above 0. Computed on the GPU. */
int oxav1_pf_superres_downscale(const uint16_t *const src[3], uint32_t width, uint32_t height, uint8_t ssx,
                                uint8_t ssy, uint8_t num_planes, uint32_t coded_width, uint16_t *const dst[3]);
/* upscale_recon (superres_elect.rs:152-232): §7.16 upscaling of u16
 * planes ((width + ssx) >> ssx etc.) through the superres kernel, errors
 * as oxav1_pf_upscale_frame. */
int oxav1_pf_superres_upscale_recon(const uint16_t *const planes[3], uint32_t width, uint32_t height,
                                    uint32_t upscaled_width, uint32_t mi_cols, uint8_t bit_depth, uint8_t ssx,
                                    uint8_t ssy, uint8_t num_planes, uint16_t *const out[3]);

/* Film-grain election helpers (encoder/film_grain_elect.rs). Planes are
 * host u16 arrays, row-major without padding; chroma planes of the
 * denoiser and the noise estimate are (width >> ssx) x (height >> ssy),
 * those of the grain application ((width + ssx) >> ssx) x
 * ((height + ssy) >> ssy), as in the reference. */

/* NoiseEstimate (film_grain_elect.rs:29-45): sigma x 16 on the 8-bit scale. */
typedef struct oxav1_noise_estimate {
    uint64_t sigma16;
    uint64_t bin_sigma16[8];
    uint64_t cb_sigma16;
    uint64_t cb_bin_sigma16[4];
    uint64_t cr_sigma16;
    uint64_t cr_bin_sigma16[4];
} oxav1_noise_estimate;

/* denoise_frame (film_grain_elect.rs:52-94): 3x3 binomial with edge
 * replication on every plane. Computed on the GPU. */
int oxav1_pf_fg_denoise(const uint16_t *const src[3], uint32_t width, uint32_t height, uint8_t ssx, uint8_t ssy,
                        uint8_t num_planes, uint16_t *const dst[3]);
/* noise_estimate (film_grain_elect.rs:97-167): the squared-difference
 * sums per intensity bin are reduced on the GPU, the square roots taken
 * on the host in double precision as the reference takes them. */
int oxav1_pf_fg_noise_estimate(const uint16_t *const input[3], const uint16_t *const denoised[3], uint32_t width,
                               uint32_t height, uint8_t ssx, uint8_t ssy, uint8_t num_planes, uint8_t bit_depth,
                               oxav1_noise_estimate *out);
/* plane_sigma16 (film_grain_elect.rs:896-910) of two n-sample planes. */
int oxav1_pf_fg_plane_sigma16(const uint16_t *a, const uint16_t *b, uint64_t n, uint8_t bit_depth,
                              uint64_t *sigma16);
/* apply_grain_to_recon (film_grain_elect.rs:834-893): §7.18.3 synthesis
 * of the u16 reconstruction into `out` through the film-grain kernels. */
int oxav1_pf_fg_apply_grain(const oxav1_film_grain_params *fg, const uint16_t *const planes[3], uint32_t width,
                            uint32_t height, uint8_t bit_depth, uint8_t ssx, uint8_t ssy, uint8_t num_planes,
                            uint8_t matrix_coefficients, uint16_t *const out[3]);

/* Loop-restoration election (encoder/lr_elect.rs). */

/* LrElectInput (encoder/lr_elect.rs:95-131). Planes are host u16 arrays,
 * row-major without padding: luma width x height, chroma chroma_w x
 * chroma_h, which must be (width + subsampling_x) >> subsampling_x and
 * (height + subsampling_y) >> subsampling_y. input is the source the
 * filters are fitted to, curr the pre-CDEF reconstruction (§7.17 reads it
 * across stripe boundaries), cdef the post-CDEF reconstruction (the
 * restoration input). restoration_type_cdf is the frame-start
 * RestorationType CDF the unit symbols are priced with ({9413, 22581,
 * 32768, 0} by default, cdf.rs:3703): non-decreasing, first entry > 0, third
 * 32768. disable_cdf_update, use_superres and superres_denom are carried
 * for the caller's write-side window (lr_elect.rs:790-806); they do not
 * change the election (every unit is priced from a fresh copy of the CDF). */
typedef struct oxav1_lr_elect_input {
    const uint16_t *input[3];
    const uint16_t *curr[3];
    const uint16_t *cdef[3];
    uint32_t width, height, chroma_w, chroma_h;
    uint8_t bit_depth, subsampling_x, subsampling_y, num_planes;
    uint32_t mi_rows, mi_cols;
    uint64_t lambda;
    uint16_t restoration_type_cdf[4];
    uint8_t disable_cdf_update, use_superres;
    uint8_t pad[2];
    uint32_t superres_denom;
} oxav1_lr_elect_input;

/* One unit's candidates as the election scored them (lr_elect.rs:444-700):
 * the SSD against the source of the unfiltered unit, of the fitted Wiener
 * filter and of each self-guided set with its fitted projection weights. */
typedef struct oxav1_lr_unit_cands {
    uint64_t d_none, d_wiener;
    uint64_t d_sgr[16];
    int16_t wiener[2][3];
    int16_t sgr_xqd[16][2];
} oxav1_lr_unit_cands;

/* LrPlan (lr_elect.rs:71-93) with the header block of header_shape
 * (lr_elect.rs:196-214): 64x64 units on every plane. elected = 0 is the
 * reference's None (every unit elected RESTORE_NONE). units[p] and
 * cands[p] are caller-owned arrays of unit_rows[p] x unit_cols[p] entries
 * (raster order), either may be NULL; units receive every unit's elected
 * filter (RESTORE_NONE units included), cands the per-unit candidate
 * records. d is the whole-frame SSD after the plan, d_pre before it; both
 * are filled whether or not a plan is elected. */
typedef struct oxav1_lr_election {
    int32_t elected;
    uint8_t frame_restoration_type[3];
    uint8_t uses_lr, uses_chroma_lr;
    uint8_t pad[3];
    uint32_t loop_restoration_size[3];
    uint32_t unit_rows[3], unit_cols[3];
    uint64_t d, d_pre;
    oxav1_lr_unit *units[3];
    oxav1_lr_unit_cands *cands[3];
} oxav1_lr_election;

/* Replaces elect_lr (encoder/lr_elect.rs:421): per plane and 64x64 unit,
 * the alternating least-squares Wiener fit, the self-guided probe fits of
 * the 16 parameter sets, the exact §7.17 evaluation of every candidate and
 * the D*256 + lambda*R256 choice with the recentred-subexp rate model, in
 * the reference's order and with its tie-breaking. The fits run on the GPU
 * in the reference's double-precision operation order (one sequential sum
 * per normal-equation term); the candidate evaluations are the LR kernels
 * reducing per-unit SSD and probe sums instead of storing samples.
 * Synchronous. */
int oxav1_pf_lr_elect(const oxav1_lr_elect_input *in, oxav1_lr_election *out);

/* apply_lr_plan (lr_elect.rs:829-912): one §7.17 run of the plan's units
 * over recon (the post-CDEF planes, restored in place) with in->curr as
 * the pre-CDEF planes, on the GPU. *ssd receives the whole-frame SSD of
 * the restored planes against in->input (in->cdef is not read). */
int oxav1_pf_lr_apply_plan(const oxav1_lr_election *plan, const oxav1_lr_elect_input *in, uint16_t *const recon[3],
                           uint64_t *ssd);

#ifdef __cplusplus
}
#endif
#endif /* OXAV1_PF_H */
