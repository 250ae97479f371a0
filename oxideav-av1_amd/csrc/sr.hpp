// §7.16 superres on gfx950: 8-tap polyphase horizontal upscale per row.
//
// Reference: superres.rs:282-445.  Each output sample is independent.  The
// per-plane constants (stepX, initialSubpelX, maxX) are derived on the host
// with the reference's 64-bit arithmetic and truncating division
// (superres.rs:363-382) and travel in the FrameJob.  A workgroup stages one
// input row and the 64x8 filter table in LDS and produces the whole output
// row from there; lanes write consecutive output samples.
//
// In the chain the deblocked frame is upscaled only on the rows loop
// restoration reads from it: the two rows above and below every stripe
// (get_source_sample, loop_restoration.rs:862-883), i.e. rows r with
// (r + (8 >> sy) + 2) mod (64 >> sy) in {0, 1, 2, 3}.
#pragma once

#include "common.hpp"

namespace oxav1 {

constexpr int kSrMaxCols = 4096;  // widest input row staged in LDS

__constant__ int16_t c_upscale_filter[64][8];  // filled from kUpscaleFilter at init

template <typename L>
struct SrSmem {
    L row[kSrMaxCols];
    int16_t filt[64 * 8];
};

// upscale_sample (superres.rs:422-445) over a row held in `row` (LDS or
// global), valid indices [0, last].
template <typename R>
__device__ __forceinline__ int32_t sr_sample(const R *row, int32_t last, int32_t x, int32_t step, int32_t init,
                                             int32_t maxx, int32_t maxv, const int16_t *filt) {
    const int32_t src_x = -(1 << 14) + init + x * step;
    const int32_t px = src_x >> 14;
    const int32_t ph = (src_x & ((1 << 14) - 1)) >> 8;
    const int16_t *f = filt + ph * 8;
    int32_t sum = 0;
    if (px - 3 >= 0 && px + 4 <= last) {
        const R *r = row + px - 3;
#pragma unroll
        for (int k = 0; k < 8; k++) sum += (int32_t)r[k] * (int32_t)f[k];
    } else {
#pragma unroll
        for (int k = 0; k < 8; k++) {
            // Clip3(0, maxX) then the row-length clamp (superres.rs:434-441)
            const int32_t sxp = clip3(0, last, clip3(0, maxx, px + (k - 3)));
            sum += (int32_t)row[sxp] * (int32_t)f[k];
        }
    }
    return clip3(0, maxv, (sum + 64) >> 7);
}

// Upscale the listed rows of one plane.  Rows are y0 + i*ystride, i < nrows,
// limited to y < ylim.
template <typename T>
__device__ void sr_rows(const T *__restrict__ in, uint32_t ipitch, int32_t in_cols, T *__restrict__ out,
                        uint32_t opitch, int32_t out_cols, int32_t y0, int32_t nrows, int32_t ylim, int32_t step,
                        int32_t init, int32_t maxx, int32_t bd, SrSmem<typename LdsOf<T>::type> &S) {
    using L = typename LdsOf<T>::type;
    const int32_t maxv = (1 << bd) - 1;
    // taps only reach columns <= min(maxX, in_cols - 1)
    const int32_t last = min(maxx, in_cols - 1);
    for (int i = threadIdx.x; i < 64 * 8; i += blockDim.x) S.filt[i] = c_upscale_filter[i >> 3][i & 7];
    const bool staged = last < kSrMaxCols;
    for (int32_t r = 0; r < nrows; r++) {
        const int32_t y = y0 + r;
        if (y < 0 || y >= ylim) continue;
        const T *src = in + (size_t)y * ipitch;
        T *orow = out + (size_t)y * opitch;
        __syncthreads();
        if (staged) {
            for (int32_t x = threadIdx.x; x <= last; x += blockDim.x) S.row[x] = (L)src[x];
            __syncthreads();
            for (int32_t x = threadIdx.x; x < out_cols; x += blockDim.x)
                orow[x] = (T)sr_sample<L>(S.row, last, x, step, init, maxx, maxv, S.filt);
        } else {
            __syncthreads();
            for (int32_t x = threadIdx.x; x < out_cols; x += blockDim.x)
                orow[x] = (T)sr_sample<T>(src, last, x, step, init, maxx, maxv, S.filt);
        }
    }
}

}  // namespace oxav1
