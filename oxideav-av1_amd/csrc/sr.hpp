// §7.16 superres on gfx950: 8-tap polyphase horizontal upscale per row.
//
// Reference: superres.rs:282-445.  Each output sample is independent.  The
// per-plane constants (stepX, initialSubpelX, maxX) are derived on the host
// with the reference's 64-bit arithmetic and truncating division
// (superres.rs:363-382) and travel in the FrameJob.  A workgroup upscales
// a run of rows; each lane produces consecutive output samples of a row so
// the 8 taps of neighbouring lanes hit the same cache lines.
#pragma once

#include "common.hpp"

namespace oxav1 {

__constant__ int16_t c_upscale_filter[64][8];  // filled from kUpscaleFilter at init

// upscale_sample (superres.rs:422-445) for output column x of one row.
template <typename T>
__device__ __forceinline__ int32_t sr_sample(const T *__restrict__ row, int32_t in_cols, int32_t x,
                                             int32_t step, int32_t init, int32_t maxx, int32_t maxv) {
    const int32_t src_x = -(1 << 14) + init + x * step;
    const int32_t px = src_x >> 14;
    const int32_t ph = (src_x & ((1 << 14) - 1)) >> 8;
    int32_t sum = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        int32_t sxp = clip3(0, maxx, px + (k - 3));
        sxp = clip3(0, in_cols - 1, sxp);
        sum += (int32_t)row[sxp] * (int32_t)c_upscale_filter[ph][k];
    }
    return clip3(0, maxv, (sum + 64) >> 7);
}

// Upscale rows [y0, y1) of one plane.
template <typename T>
__device__ void sr_rows(const T *__restrict__ in, uint32_t ipitch, int32_t in_cols, T *__restrict__ out,
                        uint32_t opitch, int32_t out_cols, int32_t y0, int32_t y1, int32_t step, int32_t init,
                        int32_t maxx, int32_t bd) {
    const int32_t maxv = (1 << bd) - 1;
    const int32_t w = out_cols;
    for (int32_t y = y0; y < y1; y++) {
        const T *row = in + (size_t)y * ipitch;
        T *orow = out + (size_t)y * opitch;
        for (int32_t x = threadIdx.x; x < w; x += blockDim.x)
            orow[x] = (T)sr_sample<T>(row, in_cols, x, step, init, maxx, maxv);
    }
}

}  // namespace oxav1
