// §7.16 superres on gfx950: 8-tap polyphase horizontal upscale per row.
//
// Reference: superres.rs:282-445.  Each output sample is independent.  The
// per-plane constants (stepX, initialSubpelX, maxX) are derived on the host
// with the reference's 64-bit arithmetic and truncating division
// (superres.rs:363-382) and travel in the FrameJob.
//
// A lane owns one output column for every row of the item: the source
// position, phase and eight taps depend on the column only, so they are
// computed once and the rows stream through with independent global loads
// (no LDS staging, no barriers).  Consecutive lanes read overlapping,
// nearly consecutive input samples and write consecutive output samples.
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
template <typename T>
__device__ __forceinline__ void sr_rows(const G<T> *__restrict__ in, uint32_t ipitch, int32_t in_cols, G<T> *__restrict__ out,
                        uint32_t opitch, int32_t out_cols, int32_t y0, int32_t nrows, int32_t ylim, int32_t step,
                        int32_t init, int32_t maxx, int32_t bd) {
    const int32_t maxv = (1 << bd) - 1;
    const int32_t last = min(maxx, in_cols - 1);  // taps never reach past this column
    const int32_t ya = max(y0, 0), yb = min(y0 + nrows, ylim);
    if (ya >= yb) return;
    for (int32_t x = threadIdx.x; x < out_cols; x += blockDim.x) {
        const int32_t src_x = -(1 << 14) + init + x * step;
        const int32_t px = src_x >> 14;
        const int32_t ph = (src_x & ((1 << 14) - 1)) >> 8;
        int32_t f[8], col[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            f[k] = c_upscale_filter[ph][k];
            col[k] = clip3(0, last, clip3(0, maxx, px + (k - 3)));
        }
        // rows in groups of four: the 32 tap loads of a group are
        // unconditional (rows clamped into the item) so they overlap
        G<T> *rout = out + (size_t)ya * opitch + x;
        for (int32_t y = ya; y < yb; y += 4, rout += 4 * opitch) {
            int32_t v[4][8];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const G<T> *row = in + (size_t)min(y + r, yb - 1) * ipitch;
#pragma unroll
                for (int k = 0; k < 8; k++) v[r][k] = (int32_t)row[col[k]];
            }
#pragma unroll
            for (int r = 0; r < 4; r++) {
                if (y + r >= yb) break;
                int32_t sum = 0;
#pragma unroll
                for (int k = 0; k < 8; k++) sum += v[r][k] * f[k];
                rout[(size_t)r * opitch] = (T)clip3(0, maxv, (sum + 64) >> 7);
            }
        }
    }
}

}  // namespace oxav1
