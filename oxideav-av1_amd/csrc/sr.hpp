// §7.16 superres on gfx950: 8-tap polyphase horizontal upscale per row.
//
// Reference: superres.rs:282-445.  Each output sample is independent.  The
// per-plane constants (stepX, initialSubpelX, maxX) are derived on the host
// with the reference's 64-bit arithmetic and truncating division
// (superres.rs:363-382) and travel in the FrameJob.
//
// A lane owns one output column for every row of the item: the source
// position, phase and eight taps depend on the column only, so they are
// computed once and the rows stream through (no LDS staging, no barriers).
//
// In the chain the deblocked frame is upscaled only on the rows loop
// restoration reads from it: the two rows above and below every stripe
// (get_source_sample, loop_restoration.rs:862-883), i.e. rows r with
// (r + (8 >> sy) + 2) mod (64 >> sy) in {0, 1, 2, 3}.
#pragma once

#include "common.hpp"

namespace oxav1 {

__constant__ int16_t c_upscale_filter[64][8];  // filled from kUpscaleFilter at init

// Upscale rows y0 .. y0+nrows-1 (limited to [0, ylim)) of one plane.
// upscale_sample (superres.rs:422-445): source position
// -(1<<14) + initialSubpelX + x*stepX, integer part px, phase from bits 8..13,
// taps at Clip3(0, maxX, px+k-3) further clamped to the stored row.
//
// A wave owns 64 consecutive output columns at a time (a segment) and walks
// its segments across the row range.  The eight taps of a segment's columns
// span at most 64 consecutive input samples (the upscale ratio is at least
// 9/8), so each lane loads one input sample per row and the taps are
// gathered across the wave with ds_bpermute: one global load per output
// sample instead of eight.  Rows go four at a time, the next four (or the
// next segment's first four) loaded before the current ones are filtered,
// so a wave waits for memory once per item, not once per segment.  The
// filter phases come from the workgroup's LDS copy of the table (ftab).  A
// segment whose taps span more than 64 samples (only very narrow planes,
// where the rounded downscaled width lifts the ratio) loads every tap.
template <typename T>
__device__ __forceinline__ void sr_rows(const G<T> *__restrict__ in, uint32_t ipitch, int32_t in_cols, G<T> *__restrict__ out,
                        uint32_t opitch, int32_t out_cols, int32_t y0, int32_t nrows, int32_t ylim, int32_t step,
                        int32_t init, int32_t maxx, int32_t bd, const int16_t *ftab) {
    const int32_t maxv = (1 << bd) - 1;
    const int32_t last = min(maxx, in_cols - 1);  // taps never reach past this column
    const int32_t ya = max(y0, 0), yb = min(y0 + nrows, ylim);
    if (ya >= yb) return;
    const int lane = threadIdx.x & 63;
    const int32_t xstep = (int32_t)blockDim.x;
    int32_t xb = (int32_t)(threadIdx.x & ~63u);
    if (xb >= out_cols) return;
    // a segment's source position, phase and (wave-uniform) first and last
    // tap columns; lanes past out_cols gather like the others and store nothing
    auto seg = [&](int32_t xs, int32_t &px, int32_t &ph, int32_t &base, int32_t &top) {
        const int32_t src_x = -(1 << 14) + init + (xs + lane) * step;
        px = src_x >> 14;
        ph = (src_x & ((1 << 14) - 1)) >> 8;
        base = __builtin_amdgcn_readfirstlane(clip3(0, last, clip3(0, maxx, px - 3)));
        top = __builtin_amdgcn_readlane(clip3(0, last, clip3(0, maxx, px + 4)), 63);
    };
    // four rows from y (clamped into the item, so every load is valid)
    auto load4 = [&](int32_t base, int32_t y, int32_t *v) {
        const uint32_t lc = (uint32_t)min(base + lane, last) * (uint32_t)sizeof(T);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const G<uint8_t> *row = (const G<uint8_t> *)(in + (size_t)min(y + r, yb - 1) * ipitch);
            v[r] = (int32_t)*(const G<T> *)(row + lc);
        }
    };
    // sample x tap + sum: u8/u16 samples are below 2^16 (said to the
    // compiler, which then forms full-rate 24-bit multiply-adds with the
    // filter's int16 taps instead of 32-bit multiplies); int32 samples keep
    // 32-bit multiplies
    auto tap = [](int32_t v, int32_t f, int32_t acc) -> int32_t {
        if constexpr (sizeof(T) <= 2) __builtin_assume((uint32_t)v < 65536u);
        return v * f + acc;
    };
    int32_t px, ph, base, top;
    seg(xb, px, ph, base, top);
    int32_t v[4] = {0, 0, 0, 0};
    if (top - base < 64) load4(base, ya, v);
    for (;;) {
        const int32_t x = xb + lane;
        const bool store = x < out_cols;
        const int32_t xn = xb + xstep;
        const bool has_next = xn < out_cols;
        int32_t pxn = 0, phn = 0, basen = 0, topn = 0;
        if (has_next) seg(xn, pxn, phn, basen, topn);
        const bool fast_next = has_next && topn - basen < 64;
        int32_t col[8];
#pragma unroll
        for (int k = 0; k < 8; k++) col[k] = clip3(0, last, clip3(0, maxx, px + (k - 3)));
        int32_t f[8];
#pragma unroll
        for (int k = 0; k < 8; k++) f[k] = ftab[ph * 8 + k];
        if (top - base < 64) {
            int32_t sel[8];
#pragma unroll
            for (int k = 0; k < 8; k++) sel[k] = (col[k] - base) << 2;
            for (int32_t y = ya; y < yb; y += 4) {
                int32_t vn[4] = {0, 0, 0, 0};
                if (y + 4 < yb)
                    load4(base, y + 4, vn);
                else if (fast_next)
                    load4(basen, ya, vn);
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    if (y + r >= yb) break;
                    int32_t sum = 0;
#pragma unroll
                    for (int k = 0; k < 8; k++) sum = tap(__builtin_amdgcn_ds_bpermute(sel[k], v[r]), f[k], sum);
                    if (store) {
                        G<uint8_t> *orow = (G<uint8_t> *)(out + (size_t)(y + r) * opitch);
                        *(G<T> *)(orow + (uint32_t)x * (uint32_t)sizeof(T)) = (T)clamp0_s((sum + 64) >> 7, maxv);
                    }
                }
#pragma unroll
                for (int r = 0; r < 4; r++) v[r] = vn[r];
            }
        } else {
            for (int32_t y = ya; y < yb; y++) {
                const G<T> *row = in + (size_t)y * ipitch;
                int32_t sum = 0;
#pragma unroll
                for (int k = 0; k < 8; k++) sum = tap((int32_t)row[col[k]], f[k], sum);
                if (store) out[(size_t)y * opitch + x] = (T)clamp0_s((sum + 64) >> 7, maxv);
            }
            if (fast_next) load4(basen, ya, v);
        }
        if (!has_next) break;
        xb = xn;
        px = pxn;
        ph = phn;
        base = basen;
        top = topn;
    }
}

}  // namespace oxav1
