// Shared device-side definitions for the gfx950 post-filter kernels.
//
// FrameJob is the device-resident description of one frame moving through
// the chain: geometry, per-stage parameters, side-information pointers and
// the HBM buffers of every stage.  The host builds it (engine.hip) and the
// stage kernels read it from a device table, so one kernel launch can cover
// many frames of different size, bit depth and subsampling.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "../../include/oxav1_pf.h"

namespace oxav1 {

// Sample storage types: u8 (8-bit), u16 (10/12-bit) for the engine, i32 for
// the drop-in entry points (the reference's PlaneBuffer samples).
enum SampleKind : uint8_t { kU8 = 0, kU16 = 1, kI32 = 2 };

template <typename T>
struct LdsOf {
    using type = uint16_t;
};
template <>
struct LdsOf<int32_t> {
    using type = int32_t;
};

struct FgScratch;  // film-grain templates / LUT / offsets (fg.hpp)

struct FrameJob {
    // ---- geometry ----
    uint32_t fw, fh, uw, mi_rows, mi_cols;
    uint8_t bd, sx, sy, np;
    uint8_t mc, use_sr, apply_lf, apply_cdef;
    uint8_t uses_lr, apply_fg, kind, pad0;
    // per-plane extents: padded (mi grid), LR input, cropped output
    uint32_t pad_rows[3], pad_cols[3];
    uint32_t lr_rows[3], lr_cols[3];
    uint32_t out_rows[3], out_cols[3];

    // ---- deblocking ----
    uint8_t lf_level[4];
    uint8_t lf_sharp, lf_delta_enabled, lf_delta_multi, pad1;
    int8_t lf_ref_deltas[8];
    int8_t lf_mode_deltas[2];
    uint8_t pad2[2];
    uint8_t seg_active[8][8];
    int16_t seg_data[8][8];
    const uint32_t *lf_mi;      // packed per-mi words (OXAV1_MI_* layout)
    const int8_t *delta_lf;     // 4 per mi or null
    const uint8_t *tx[3];       // per-plane loop-filter tx grid
    uint32_t tx_cols[3];

    // ---- CDEF ----
    oxav1_cdef_params cdef;
    const int8_t *cdef_idx;     // per 64x64 unit
    uint32_t cdef_ucols, cdef_urows;
    uint32_t cdef_ext_rows[3], cdef_ext_cols[3];  // extent the 8x8 walk writes

    // ---- superres (per plane, computed on the host in 64-bit) ----
    int32_t sr_step[3], sr_init[3], sr_maxx[3];
    uint32_t sr_in_w[3];        // downscaled plane width (input columns used)

    // ---- loop restoration ----
    union {
        uint8_t lr_type[4];     // per plane; [3] unused
        uint32_t lr_type4;      // the four bytes as one aligned word
    };
    uint32_t lr_unit_size[3], lr_unit_rows[3], lr_unit_cols[3];
    const oxav1_lr_unit *lr_units[3];
    // multiply-shift reciprocals (lr_mdiv) of lr_tiles_x and lr_unit_size
    uint64_t lr_magic_tx[3], lr_magic_us[3];
    uint32_t lr_tile_w[3];      // LR tile width (divides the unit size)
    uint32_t lrc_rows[3], lrc_cols[3];  // extents of the LR "curr" buffer
    uint8_t lr_copy_none;       // tiles without restoration copy CDEF -> out
    uint8_t pad4[3];

    // ---- film grain ----
    oxav1_film_grain_params fg;
    FgScratch *fgs;

    // ---- buffers (pitches in samples) ----
    const void *src[3];
    uint32_t src_pitch[3];
    void *deb[3];               // deblocked (padded)
    uint32_t deb_pitch[3];
    void *cdf[3];               // CDEF output (padded)
    uint32_t cdf_pitch[3];
    void *up_cdef[3];           // LR "cdef" input (upscaled or == cdf)
    uint32_t up_cdef_pitch[3];
    void *up_deb[3];            // LR "curr" input (upscaled or == deb)
    uint32_t up_deb_pitch[3];
    void *ref_out[3];           // cropped pre-grain frame
    uint32_t ref_pitch[3];
    void *out[3];               // film-grain output
    uint32_t out_pitch[3];

    // ---- work items per stage ----
    // stage s, plane p: items [item_base[s][p], item_base[s][p+1]) of this frame;
    // for superres planes 3..5 are the deblocked set.
    uint32_t item_base[5][7];
    uint32_t lf_tiles_x[3];
    uint32_t lr_tiles_x[3];
    uint32_t fg_tiles_x[3];
    uint32_t sr_rows_per_item;
    uint8_t lf_plane_on[3];     // plane filtered (loop_filter.rs:276-280)
    uint8_t sr_deb_on;          // upscale the deblocked set too
};

enum Stage : int { kStageLf = 0, kStageCdef = 1, kStageSr = 2, kStageLr = 3, kStageFg = 4 };

__device__ __forceinline__ int32_t clip3(int32_t a, int32_t b, int32_t x) {
    return x < a ? a : (x > b ? b : x);
}
// Median of three signed integers in one VALU instruction (v_med3_i32); with
// lo <= hi, med3_i32(v, lo, hi) is v clamped to [lo, hi].
__device__ __forceinline__ int32_t med3_i32(int32_t a, int32_t b, int32_t c) {
    int32_t r;
    asm("v_med3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// The same with wave-uniform bounds: hi from a scalar register (one scalar
// operand is all a VOP3 instruction reads), lo 0 as an inline constant.
__device__ __forceinline__ int32_t med3_i32_s(int32_t v, int32_t lo, int32_t hi) {
    int32_t r;
    asm("v_med3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(v), "v"(lo), "s"(hi));
    return r;
}
__device__ __forceinline__ int32_t clamp0_s(int32_t v, int32_t hi) {
    int32_t r;
    asm("v_med3_i32 %0, %1, 0, %2" : "=v"(r) : "v"(v), "s"(hi));
    return r;
}

// Products of operands that fit 24 bits (two's complement for the i24 forms,
// unsigned for the u24 forms) as single full-rate instructions: the low 32
// bits of x * y (+ z).  The compiler picks these by itself only where it can
// derive both operand ranges; for sums of many samples, products of products
// and table values it falls back to the quarter-rate v_mul_lo_u32.
__device__ __forceinline__ int32_t mul_i24(int32_t x, int32_t y) {
    int32_t r;
    asm("v_mul_i32_i24 %0, %1, %2" : "=v"(r) : "v"(x), "v"(y));
    return r;
}
__device__ __forceinline__ int32_t mad_i24(int32_t x, int32_t y, int32_t z) {
    int32_t r;
    asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(r) : "v"(x), "v"(y), "v"(z));
    return r;
}
__device__ __forceinline__ uint32_t mul_u24(uint32_t x, uint32_t y) {
    uint32_t r;
    asm("v_mul_u32_u24 %0, %1, %2" : "=v"(r) : "v"(x), "v"(y));
    return r;
}
// The same with the first factor k wave-uniform, read from a scalar
// register (k must not vary across the wave: a varying value would be
// replaced by its first active lane's).
__device__ __forceinline__ int32_t mad_i24_s(int32_t k, int32_t x, int32_t z) {
    int32_t r;
    asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(r) : "s"(k), "v"(x), "v"(z));
    return r;
}
__device__ __forceinline__ uint32_t mul_u24_s(uint32_t k, uint32_t x) {
    uint32_t r;
    asm("v_mul_u32_u24 %0, %1, %2" : "=v"(r) : "s"(k), "v"(x));
    return r;
}
// x * K for |x| < 2^23 and a small constant K (the instruction's inline or
// literal operand): the compiler keeps multiplications by small constants
// such as 3 on the quarter-rate 32-bit multiplier.
template <int K>
__device__ __forceinline__ int32_t mul_i24k(int32_t x) {
    static_assert(K >= -(1 << 23) && K < (1 << 23), "24-bit constant");
    int32_t r;
    asm("v_mul_i32_i24 %0, %1, %2" : "=v"(r) : "n"(K), "v"(x));
    return r;
}
// x * K for a constant K < 2^24, K as the instruction's literal operand.
template <uint32_t K>
__device__ __forceinline__ uint32_t mul_u24k(uint32_t x) {
    static_assert(K < (1u << 24), "24-bit constant");
    uint32_t r;
    asm("v_mul_u32_u24 %0, %1, %2" : "=v"(r) : "n"(K), "v"(x));
    return r;
}

__device__ __forceinline__ int32_t round2(int32_t x, uint32_t n) {
    return n == 0 ? x : (x + (1 << (n - 1))) >> n;
}
__device__ __forceinline__ uint32_t floor_log2(uint32_t x) {
    return x == 0 ? 0u : 31u - (uint32_t)__clz((int)x);
}

// Pointers in the FrameJob are generic (flat) addresses; the frame buffers
// all live in device global memory, so casting them to address space 1 lets
// the compiler emit global_load/global_store instead of flat accesses (which
// also tie up the LDS counter and cannot use the scalar-offset forms).
template <typename T>
using G = __attribute__((address_space(1))) T;
// LDS (address space 3), for helpers that are not inlined into the kernel.
template <typename T>
using S3 = __attribute__((address_space(3))) T;

// Element i of a side-information array the FrameJob points at, read as a
// global load.
template <typename T>
__device__ __forceinline__ T gread(const T *p, size_t i) {
    return ((const G<T> *)p)[i];
}

// A value every lane of the wave holds, moved to scalar registers.  FrameJob
// fields arrive in vector registers (they are read with vector loads), so
// without this the compiler forms every address derived from them per lane in
// 64-bit vector arithmetic instead of a scalar base plus a 32-bit offset.
__device__ __forceinline__ uint32_t wave_u32(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
template <typename T>
__device__ __forceinline__ T *wave_ptr(T *p) {
    const uint64_t v = (uint64_t)p;
    return (T *)(((uint64_t)wave_u32((uint32_t)(v >> 32)) << 32) | wave_u32((uint32_t)v));
}

template <typename T>
__device__ __forceinline__ G<T> *plane_ptr(void *p) {
    return (G<T> *)p;
}
template <typename T>
__device__ __forceinline__ const G<T> *plane_ptr(const void *p) {
    return (const G<T> *)p;
}

}  // namespace oxav1
