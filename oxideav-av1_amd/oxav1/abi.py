"""ctypes view of the C ABI in include/oxav1_pf.h and include/oxav1_gen.h.

The struct layouts mirror the C declarations field for field; every struct
is checked against the library's own sizeof (oxav1_pf_struct_size) when the
library is loaded.  The shared library is built in-tree
(oxideav-av1_amd/liboxav1_pf.so); loading fails loudly if it is missing.
"""
import ctypes as C
import os

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "liboxav1_pf.so")

OK = 0
ERR_UPSCALED_NOT_LARGER = 1
ERR_PLANE_SHAPE_MISMATCH = 2
ERR_PLANE_COUNT_MISMATCH = 3
ERR_INVALID_ARG = 10
ERR_ALIAS = 11
ERR_HIP = 12
ERR_OUT_OF_MEMORY = 13
ERR_ADMISSION_REFUSED = 14
ERR_UNSUPPORTED = 15
ERR_NO_DEVICE = 16
ERR_BUSY = 17
ERR_SCRATCH_TOO_SMALL = 18

RESTORE_NONE, RESTORE_WIENER, RESTORE_SGRPROJ, RESTORE_SWITCHABLE = 0, 1, 2, 3
MI_SIZE_SHIFT, MI_SKIP_BIT, MI_INTRA_BIT, MI_REF_SHIFT, MI_MODE_SHIFT, MI_SEG_SHIFT = 0, 5, 6, 8, 12, 24

u8, i8, u16, i16, u32, i32, u64 = (C.c_uint8, C.c_int8, C.c_uint16, C.c_int16, C.c_uint32,
                                   C.c_int32, C.c_uint64)


class Plane(C.Structure):
    _fields_ = [("samples", C.POINTER(i32)), ("rows", u32), ("cols", u32)]


class LfCtx(C.Structure):
    _fields_ = [
        ("loop_filter_level", u8 * 4), ("loop_filter_sharpness", u8),
        ("loop_filter_delta_enabled", u8), ("delta_lf_multi", u8), ("num_planes", u8),
        ("loop_filter_ref_deltas", i8 * 8), ("loop_filter_mode_deltas", i8 * 2),
        ("bit_depth", u8), ("subsampling_x", u8), ("subsampling_y", u8), ("reserved0", u8),
        ("mi_rows", u32), ("mi_cols", u32), ("frame_width", u32), ("frame_height", u32),
    ]


class LfGrids(C.Structure):
    _fields_ = [
        ("is_intra", C.POINTER(u8)), ("skip", C.POINTER(u8)), ("ref_frame", C.POINTER(u8)),
        ("mode", C.POINTER(u8)), ("segment_id", C.POINTER(u8)), ("delta_lf", C.POINTER(i8)),
        ("mi_size", C.POINTER(u8)), ("lf_tx_size", C.POINTER(u8) * 3),
        ("seg_feature_active", (u8 * 8) * 8), ("seg_feature_data", (i16 * 8) * 8),
    ]


class CdefParams(C.Structure):
    _fields_ = [
        ("cdef_damping", u8), ("cdef_bits", u8), ("y_pri_strength", u8 * 8),
        ("y_sec_strength", u8 * 8), ("uv_pri_strength", u8 * 8), ("uv_sec_strength", u8 * 8),
    ]


class CdefCtx(C.Structure):
    _fields_ = [
        ("mi_rows", u32), ("mi_cols", u32), ("num_planes", u8), ("bit_depth", u8),
        ("subsampling_x", u8), ("subsampling_y", u8), ("params", CdefParams),
    ]


class CdefGrids(C.Structure):
    _fields_ = [("cdef_idx", C.POINTER(i8)), ("skip", C.POINTER(u8))]


class SrCtx(C.Structure):
    _fields_ = [
        ("use_superres", u8), ("num_planes", u8), ("bit_depth", u8), ("subsampling_x", u8),
        ("subsampling_y", u8), ("reserved", u8 * 3), ("frame_width", u32),
        ("upscaled_width", u32), ("frame_height", u32), ("mi_cols", u32),
    ]


class LrUnit(C.Structure):
    _fields_ = [("type", u8), ("sgr_set", u8), ("wiener", (i16 * 3) * 2), ("sgr_xqd", i16 * 2)]


class LrCtx(C.Structure):
    _fields_ = [
        ("mi_rows", u32), ("mi_cols", u32), ("num_planes", u8), ("bit_depth", u8),
        ("subsampling_x", u8), ("subsampling_y", u8), ("frame_height", u32),
        ("upscaled_width", u32), ("frame_restoration_type", u8 * 3), ("uses_lr", u8),
        ("loop_restoration_size", u32 * 3),
    ]


class LrGrids(C.Structure):
    _fields_ = [("units", C.POINTER(LrUnit) * 3)]


class FilmGrainParams(C.Structure):
    _fields_ = [
        ("apply_grain", u8), ("update_grain", u8), ("grain_seed", u16), ("num_y_points", u8),
        ("point_y_value", u8 * 14), ("point_y_scaling", u8 * 14),
        ("chroma_scaling_from_luma", u8), ("num_cb_points", u8), ("point_cb_value", u8 * 10),
        ("point_cb_scaling", u8 * 10), ("num_cr_points", u8), ("point_cr_value", u8 * 10),
        ("point_cr_scaling", u8 * 10), ("grain_scaling", u8), ("ar_coeff_lag", u8),
        ("ar_coeffs_y_plus_128", u8 * 24), ("ar_coeffs_cb_plus_128", u8 * 25),
        ("ar_coeffs_cr_plus_128", u8 * 25), ("ar_coeff_shift", u8), ("grain_scale_shift", u8),
        ("cb_mult", u8), ("cb_luma_mult", u8), ("cb_offset", u16), ("cr_mult", u8),
        ("cr_luma_mult", u8), ("cr_offset", u16), ("overlap_flag", u8),
        ("clip_to_restricted_range", u8),
    ]


class FrameDesc(C.Structure):
    _fields_ = [
        ("frame_width", u32), ("frame_height", u32), ("upscaled_width", u32), ("mi_rows", u32),
        ("mi_cols", u32), ("bit_depth", u8), ("subsampling_x", u8), ("subsampling_y", u8),
        ("num_planes", u8), ("matrix_coefficients", u8), ("use_superres", u8), ("apply_lf", u8),
        ("apply_cdef", u8), ("lf", LfCtx), ("cdef", CdefParams), ("lr", LrCtx),
        ("fg", FilmGrainParams), ("seg_feature_active", (u8 * 8) * 8),
        ("seg_feature_data", (i16 * 8) * 8),
    ]


class FrameDev(C.Structure):
    _fields_ = [
        ("src", C.c_void_p * 3), ("src_pitch", u32 * 3), ("ref_out", C.c_void_p * 3),
        ("ref_pitch", u32 * 3), ("out", C.c_void_p * 3), ("out_pitch", u32 * 3),
        ("lf_mi", C.c_void_p), ("delta_lf", C.c_void_p), ("lf_tx_size", C.c_void_p * 3),
        ("cdef_idx", C.c_void_p), ("lr_units", C.c_void_p * 3),
    ]


class FrameHost(C.Structure):
    _fields_ = [
        ("src", C.c_void_p * 3), ("src_pitch", u32 * 3), ("lf_mi", C.c_void_p), ("delta_lf", C.c_void_p),
        ("lf_tx_size", C.c_void_p * 3), ("cdef_idx", C.c_void_p), ("lr_units", C.c_void_p * 3),
        ("ref_out", C.c_void_p * 3), ("out", C.c_void_p * 3),
    ]


class GenConfig(C.Structure):
    _fields_ = [
        ("upscaled_width", u32), ("frame_height", u32), ("bit_depth", u8), ("subsampling_x", u8),
        ("subsampling_y", u8), ("mono", u8), ("sb128", u8), ("superres_denom", u8),
        ("lr_unit_shift", u8), ("lr_uv_shift", u8), ("enable_lf", u8), ("enable_cdef", u8),
        ("enable_lr", u8), ("enable_fg", u8), ("fg_overlap", u8), ("fg_num_y_points", u8),
        ("fg_ar_lag", u8), ("fg_chroma_points", u8), ("noise_amp", u8), ("rect_partitions", u8),
        ("exercise_segments", u8), ("exercise_delta_lf", u8), ("cdef_skip_units", u8),
        ("lr_types", u8), ("decoder_geometry", u8),
    ]


class GenFrame(C.Structure):
    _fields_ = [
        ("desc", FrameDesc), ("planes", C.POINTER(i32) * 3), ("plane_rows", u32 * 3),
        ("plane_cols", u32 * 3), ("is_intra", C.POINTER(u8)), ("skip", C.POINTER(u8)),
        ("ref_frame", C.POINTER(u8)), ("mode", C.POINTER(u8)), ("segment_id", C.POINTER(u8)),
        ("mi_size", C.POINTER(u8)), ("delta_lf", C.POINTER(i8)), ("lf_tx_size", C.POINTER(u8) * 3),
        ("tx_rows", u32 * 3), ("tx_cols", u32 * 3), ("lf_mi", C.POINTER(u32)),
        ("cdef_idx", C.POINTER(i8)), ("cdef_rows", u32), ("cdef_cols", u32),
        ("lr_units", C.POINTER(LrUnit) * 3), ("lr_unit_rows", u32 * 3), ("lr_unit_cols", u32 * 3),
    ]


class WalkerGrids(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("is_inter", "skip", "ref_frame", "mode", "segment_id", "mi_size",
                                           "tx_size", "inter_tx_size")]


class Tuning(C.Structure):
    _fields_ = [("blocks_per_cu", (C.c_uint16 * 2) * 5), ("waves_per_eu", (u8 * 2) * 5),
                ("sr_rows_per_item", u8 * 2), ("lr_tile_w", u8 * 2), ("sub_batches", u8 * 2),
                ("items_per_claim", (u8 * 2) * 5)]


class ServerConfig(C.Structure):
    _fields_ = [("lanes", u32), ("max_batch", u32), ("arena_bytes", u64), ("largest", C.c_void_p)]


class ServerDone(C.Structure):
    _fields_ = [
        ("ticket", u64), ("session", u32), ("slot", u32), ("status", C.c_int32), ("batch", u32),
        ("arrival_us", C.c_double), ("launch_us", C.c_double), ("done_us", C.c_double),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("sessions_open", u32), ("lanes_busy", u32), ("frames_queued", u32), ("max_batch_seen", u32),
        ("frames_submitted", u64), ("frames_done", u64), ("launches", u64), ("arena_bytes", u64),
        ("arena_used", u64), ("lane_bytes", u64),
    ]


class CdefCand(C.Structure):
    _fields_ = [("damping", u8), ("plane_set", u8), ("pri", u8), ("sec", u8)]


class CdefSsdFrame(C.Structure):
    _fields_ = [
        ("mi_rows", u32), ("mi_cols", u32), ("cdef_idx", C.c_void_p), ("skip", C.c_void_p),
        ("recon", C.c_void_p * 3), ("input", C.c_void_p * 3), ("recon_pitch", u32 * 3),
        ("input_pitch", u32 * 3), ("rows", u32 * 3), ("cols", u32 * 3), ("bit_depth", u8),
        ("subsampling_x", u8), ("subsampling_y", u8), ("num_planes", u8), ("sample_bytes", u8),
        ("pad", u8 * 3), ("ssd", C.c_void_p),
    ]


class CdefElectInput(C.Structure):
    _fields_ = [
        ("mi_rows", u32), ("mi_cols", u32), ("cdef_idx", C.c_void_p), ("skip", C.c_void_p),
        ("input", C.c_void_p * 3), ("recon", C.c_void_p * 3), ("width", u32), ("height", u32),
        ("chroma_w", u32), ("chroma_h", u32), ("bit_depth", u8), ("subsampling_x", u8),
        ("subsampling_y", u8), ("num_planes", u8), ("max_bits", u8), ("pad", u8 * 3), ("lambda_", u64),
    ]


class CdefElection(C.Structure):
    _fields_ = [
        ("elected", C.c_int32), ("has_frame_level", C.c_int32), ("best", CdefParams),
        ("frame_level", CdefParams), ("best_d", u64), ("frame_level_d", u64), ("base_d", u64),
        ("best_unit_idx", C.c_void_p), ("frame_level_unit_idx", C.c_void_p),
    ]


class LrElectInput(C.Structure):
    _fields_ = [
        ("input", C.c_void_p * 3), ("curr", C.c_void_p * 3), ("cdef", C.c_void_p * 3),
        ("width", u32), ("height", u32), ("chroma_w", u32), ("chroma_h", u32),
        ("bit_depth", u8), ("subsampling_x", u8), ("subsampling_y", u8), ("num_planes", u8),
        ("mi_rows", u32), ("mi_cols", u32), ("lambda_", u64), ("restoration_type_cdf", C.c_uint16 * 4),
        ("disable_cdf_update", u8), ("use_superres", u8), ("pad", u8 * 2), ("superres_denom", u32),
    ]


class LrUnitCands(C.Structure):
    _fields_ = [("d_none", u64), ("d_wiener", u64), ("d_sgr", u64 * 16), ("wiener", (i16 * 3) * 2),
                ("sgr_xqd", (i16 * 2) * 16)]


class LrElection(C.Structure):
    _fields_ = [
        ("elected", C.c_int32), ("frame_restoration_type", u8 * 3), ("uses_lr", u8), ("uses_chroma_lr", u8),
        ("pad", u8 * 3), ("loop_restoration_size", u32 * 3), ("unit_rows", u32 * 3), ("unit_cols", u32 * 3),
        ("d", u64), ("d_pre", u64), ("units", C.c_void_p * 3), ("cands", C.c_void_p * 3),
    ]


class NoiseEstimate(C.Structure):
    _fields_ = [("sigma16", u64), ("bin_sigma16", u64 * 8), ("cb_sigma16", u64), ("cb_bin_sigma16", u64 * 4),
                ("cr_sigma16", u64), ("cr_bin_sigma16", u64 * 4)]


ABI_STRUCTS = {
    "oxav1_plane_i32": Plane, "oxav1_lf_ctx": LfCtx, "oxav1_lf_grids": LfGrids,
    "oxav1_cdef_params": CdefParams, "oxav1_cdef_ctx": CdefCtx, "oxav1_cdef_grids": CdefGrids,
    "oxav1_sr_ctx": SrCtx, "oxav1_lr_unit": LrUnit, "oxav1_lr_ctx": LrCtx,
    "oxav1_lr_grids": LrGrids, "oxav1_film_grain_params": FilmGrainParams,
    "oxav1_frame_desc": FrameDesc, "oxav1_frame_dev": FrameDev, "oxav1_frame_host": FrameHost,
    "oxav1_tuning": Tuning, "oxav1_walker_grids": WalkerGrids, "oxav1_server_config": ServerConfig, "oxav1_server_done": ServerDone,
    "oxav1_server_stats": ServerStats, "oxav1_cdef_cand": CdefCand, "oxav1_cdef_ssd_frame": CdefSsdFrame,
    "oxav1_cdef_elect_input": CdefElectInput, "oxav1_cdef_election": CdefElection,
    "oxav1_noise_estimate": NoiseEstimate, "oxav1_lr_elect_input": LrElectInput,
    "oxav1_lr_unit_cands": LrUnitCands, "oxav1_lr_election": LrElection,
}

_P = C.POINTER
_SIGS = {
    "oxav1_pf_init": (C.c_int, []),
    "oxav1_pf_strerror": (C.c_char_p, [C.c_int]),
    "oxav1_pf_last_hip_error": (C.c_char_p, []),
    "oxav1_pf_table_checksum": (C.c_int64, [C.c_char_p]),
    "oxav1_pf_struct_size": (C.c_size_t, [C.c_char_p]),
    "oxav1_pf_count_units_in_frame": (u32, [u32, u32]),
    "oxav1_pf_loop_filter_frame": (C.c_int, [_P(LfCtx), _P(LfGrids), _P(Plane), u32]),
    "oxav1_pf_cdef_frame": (C.c_int, [_P(CdefCtx), _P(CdefGrids), _P(Plane), u32, _P(Plane), u32]),
    "oxav1_pf_upscale_frame": (C.c_int, [_P(SrCtx), _P(Plane), u32, _P(Plane), u32]),
    "oxav1_pf_loop_restoration_frame": (
        C.c_int, [_P(LrCtx), _P(LrGrids), _P(Plane), u32, _P(Plane), u32, _P(Plane), u32]),
    "oxav1_pf_film_grain_synthesis": (
        C.c_int, [_P(FilmGrainParams), u8, u8, u8, u8, u8, _P(Plane), u32]),
    "oxav1_pf_engine_create": (C.c_int, [_P(C.c_void_p), C.c_void_p, u32]),
    "oxav1_pf_engine_destroy": (C.c_int, [C.c_void_p]),
    "oxav1_pf_engine_prepare": (C.c_int, [C.c_void_p, _P(FrameDesc), _P(FrameDev), u32]),
    "oxav1_pf_engine_run": (C.c_int, [C.c_void_p, C.c_int]),
    "oxav1_pf_engine_run_frame": (C.c_int, [C.c_void_p, _P(FrameDesc), _P(FrameDev)]),
    "oxav1_pf_run_frame_noalloc": (C.c_int, [_P(FrameDesc), _P(FrameDev), C.c_void_p]),
    "oxav1_pf_dropin_narrowing": (C.c_int, [C.c_int]),
    "oxav1_pf_engine_run_batch": (C.c_int, [C.c_void_p, _P(FrameDesc), _P(FrameDev), u32, C.c_int]),
    "oxav1_pf_engine_sync": (C.c_int, [C.c_void_p]),
    "oxav1_pf_engine_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "oxav1_pf_engine_stage_ms": (C.c_int, [C.c_void_p, C.c_char_p, _P(C.c_double), _P(u32)]),
    "oxav1_pf_engine_stream": (C.c_void_p, [C.c_void_p]),
    "oxav1_pf_engine_query": (C.c_int, [C.c_void_p]),
    "oxav1_pf_build_grids": (C.c_int, [_P(WalkerGrids), u32, u32, u8, u8, u8, C.c_void_p, C.c_void_p * 3,
                                       C.c_void_p]),
    "oxav1_pf_engine_set_tuning": (C.c_int, [C.c_void_p, _P(Tuning)]),
    "oxav1_pf_engine_get_tuning": (C.c_int, [C.c_void_p, _P(Tuning), _P(u32), _P((u32 * 4) * 5)]),
    "oxav1_pf_engine_reserve": (C.c_int, [C.c_void_p, _P(FrameDesc), u32, _P(u64)]),
    "oxav1_pf_server_now_us": (C.c_double, []),
    "oxav1_pf_server_create": (C.c_int, [_P(C.c_void_p), _P(ServerConfig)]),
    "oxav1_pf_server_destroy": (C.c_int, [C.c_void_p]),
    "oxav1_pf_server_open_session": (
        C.c_int, [C.c_void_p, _P(FrameDesc), u32, u64, _P(u32), _P(C.c_void_p), _P(u64)]),
    "oxav1_pf_server_close_session": (C.c_int, [C.c_void_p, u32]),
    "oxav1_pf_server_slot": (C.c_int, [C.c_void_p, u32, u32, u32, C.c_int, _P(C.c_void_p), _P(u32)]),
    "oxav1_pf_server_submit": (
        C.c_int, [C.c_void_p, u32, _P(FrameDesc), _P(FrameDev), u32, C.c_double, _P(u64)]),
    "oxav1_pf_server_submit_host": (
        C.c_int, [C.c_void_p, u32, _P(FrameDesc), _P(FrameHost), u32, C.c_double, _P(u64)]),
    "oxav1_pf_server_frame_input_bytes": (u64, [_P(FrameDesc)]),
    "oxav1_pf_host_alloc": (C.c_int, [_P(C.c_void_p), C.c_size_t]),
    "oxav1_pf_host_free": (C.c_int, [C.c_void_p]),
    "oxav1_pf_server_submit_existing": (
        C.c_int, [C.c_void_p, u32, u32, _P(FilmGrainParams), C.c_double, _P(u64)]),
    "oxav1_pf_server_poll": (C.c_int, [C.c_void_p, _P(ServerDone), u32, _P(u32)]),
    "oxav1_pf_server_stats": (C.c_int, [C.c_void_p, _P(ServerStats)]),
    "oxav1_pf_cdef_unit_ssd": (C.c_int, [_P(CdefSsdFrame), u32, _P(CdefCand), u32, C.c_void_p]),
    "oxav1_pf_cdef_elect": (C.c_int, [_P(CdefElectInput), _P(CdefElection)]),
    "oxav1_pf_superres_coded_width": (u32, [u32, u32]),
    "oxav1_pf_superres_candidate_denoms": (u32, [u32, _P(u32)]),
    "oxav1_pf_superres_probe": (C.c_int, [C.c_void_p, u32, u32, u8, _P(C.c_int32)]),
    "oxav1_pf_superres_downscale": (C.c_int, [C.c_void_p * 3, u32, u32, u8, u8, u8, u32, C.c_void_p * 3]),
    "oxav1_pf_superres_upscale_recon": (
        C.c_int, [C.c_void_p * 3, u32, u32, u32, u32, u8, u8, u8, u8, C.c_void_p * 3]),
    "oxav1_pf_fg_denoise": (C.c_int, [C.c_void_p * 3, u32, u32, u8, u8, u8, C.c_void_p * 3]),
    "oxav1_pf_fg_noise_estimate": (
        C.c_int, [C.c_void_p * 3, C.c_void_p * 3, u32, u32, u8, u8, u8, u8, _P(NoiseEstimate)]),
    "oxav1_pf_fg_plane_sigma16": (C.c_int, [C.c_void_p, C.c_void_p, u64, u8, _P(u64)]),
    "oxav1_pf_fg_apply_grain": (
        C.c_int, [_P(FilmGrainParams), C.c_void_p * 3, u32, u32, u8, u8, u8, u8, u8, C.c_void_p * 3]),
    "oxav1_pf_lr_elect": (C.c_int, [_P(LrElectInput), _P(LrElection)]),
    "oxav1_pf_lr_apply_plan": (C.c_int, [_P(LrElection), _P(LrElectInput), C.c_void_p * 3, _P(u64)]),
    "oxav1_pf_dev_alloc": (C.c_int, [_P(C.c_void_p), C.c_size_t]),
    "oxav1_pf_dev_free": (C.c_int, [C.c_void_p]),
    "oxav1_pf_dev_upload": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "oxav1_pf_dev_download": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "oxav1_pf_dev_fill": (C.c_int, [C.c_void_p, C.c_int, C.c_size_t]),
    "oxav1_pf_dev_info": (C.c_int, [_P(C.c_int), _P(C.c_size_t), _P(C.c_size_t)]),
    "oxav1_pf_dev_copy_kernel": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]),
    "oxav1_gen_default_config": (C.c_int, [C.c_int, u64, _P(GenConfig)]),
    "oxav1_gen_frame_create": (C.c_int, [_P(GenConfig), u64, _P(_P(GenFrame))]),
    "oxav1_gen_frame_destroy": (None, [_P(GenFrame)]),
    "oxav1_gen_splitmix64": (u64, [u64]),
}

_lib = None


class Oxav1Error(RuntimeError):
    def __init__(self, status, what=""):
        lib = load()
        msg = lib.oxav1_pf_strerror(status).decode()
        if status == ERR_HIP:
            msg += " (" + lib.oxav1_pf_last_hip_error().decode() + ")"
        super().__init__(f"{what}: status {status}: {msg}" if what else f"status {status}: {msg}")
        self.status = status


def load(path=None):
    """Load liboxav1_pf.so (built in-tree); raises if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FileNotFoundError(
            f"{p} is missing: build it with `make -C oxideav-av1_amd` or __graft_entry__.build()")
    lib = C.CDLL(p)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    for cname, cls in ABI_STRUCTS.items():
        n = lib.oxav1_pf_struct_size(cname.encode())
        if n != C.sizeof(cls):
            raise RuntimeError(f"ABI mismatch for {cname}: C {n} bytes, ctypes {C.sizeof(cls)}")
    if path is None:
        _lib = lib
    return lib


def check(status, what=""):
    if status != OK:
        raise Oxav1Error(status, what)
    return status
