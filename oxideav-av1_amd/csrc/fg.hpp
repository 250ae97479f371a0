// §7.18.3 film-grain synthesis on gfx950.
//
// Reference: film_grain.rs:100-873.  Two kernels per frame:
//   fg_prep  (one workgroup): scaling LUTs (§7.18.3.4), the per-(stripe,
//            32x32 block) LFSR offsets, and the grain templates (§7.18.3.3).
//            The white-noise LFSR is linear over GF(2), so each lane jumps to
//            its chunk of the sequence with precomputed powers of the step
//            matrix; the causal auto-regressive filter runs as a wavefront
//            with one lane per template row, row y trailing row y-1 by four
//            columns (the filter reaches x+3 on earlier rows).
//   fg_blend (whole frame): per output sample, the noise value is rebuilt
//            from the templates and offsets, including the horizontal and
//            vertical 27/17 (or 23/22) overlap blends of the reference's
//            noise stripes (film_grain.rs:626-739), then blended with the
//            scaling LUT (film_grain.rs:759-828).  Chroma reads the
//            pre-noise luma, as the reference blends chroma before luma.
#pragma once

#include "common.hpp"

namespace oxav1 {

constexpr int kGrainW = 82, kGrainH = 73;
constexpr int kTqW = 84;  // row pitch of the shifted template copies (a multiple of four)

struct FgScratch {
    int16_t luma[kGrainH * kGrainW];
    int16_t cb[kGrainH * kGrainW];
    int16_t cr[kGrainH * kGrainW];
    uint8_t lut[3][256];
    int32_t chroma_w, chroma_h;
    uint32_t nstripes, nblocks;
    uint8_t *offsets;  // nstripes * nblocks LFSR draws
    // tq[p][k]: plane p's template shifted left by k columns (row pitch
    // kTqW, zero past the template), so the four values at any column c of
    // a row are one aligned 8-byte read from copy c & 3 at column c & ~3
    alignas(16) int16_t tq[3][4][kGrainH * kTqW];
};

__constant__ int16_t c_gauss[2048];       // filled from kGaussianSequence at init
__constant__ uint16_t c_lfsr_jump[16][16]; // [b][i]: column i of the step matrix ^ (2^b)

// RandomRegister::next (film_grain.rs:117-130)
__device__ __forceinline__ uint32_t fg_step(uint32_t &r) {
    const uint32_t bit = (r ^ (r >> 1) ^ (r >> 3) ^ (r >> 12)) & 1u;
    r = ((r >> 1) | (bit << 15)) & 0xffffu;
    return r;
}
// state after n steps
__device__ __forceinline__ uint32_t fg_jump(uint32_t r, uint32_t n) {
    for (int b = 0; b < 16; b++) {
        if (!((n >> b) & 1u)) continue;
        uint32_t v = 0;
#pragma unroll
        for (int i = 0; i < 16; i++)
            if ((r >> i) & 1u) v ^= c_lfsr_jump[b][i];
        r = v;
    }
    return r;
}

__device__ __forceinline__ int32_t fg_round2(int32_t x, uint32_t n) { return n == 0 ? x : (x + (1 << (n - 1))) >> n; }

// scaling_lookup_fill (film_grain.rs:480-526) for one table entry v: the
// left plateau, then the ramps in segment order (a later segment wins where
// non-increasing points make ramps overlap), then the right plateau;
// entries nothing covers stay 0, as in the reference's fresh table.
__device__ __forceinline__ uint8_t fg_lut_value(uint32_t np_, const uint8_t *xs, const uint8_t *ys, int32_t v) {
    if (np_ == 0) return 0;
    int32_t val = v < (int32_t)xs[0] ? (int32_t)ys[0] : 0;
    for (uint32_t i = 0; i + 1 < np_; i++) {
        const int32_t xa = xs[i], xb = xs[i + 1], ya = ys[i], yb = ys[i + 1];
        const int32_t dx = xb - xa;
        if (dx <= 0 || v < xa || v >= xb) continue;
        const int32_t delta = (yb - ya) * ((65536 + (dx >> 1)) / dx);
        val = clip3(0, 255, ya + (((v - xa) * delta + 32768) >> 16));
    }
    if (v >= (int32_t)xs[np_ - 1]) val = ys[np_ - 1];
    return (uint8_t)val;
}

// White noise for one template (generate_grain white-noise loops).
__device__ void fg_white(int16_t *g, int count, uint32_t seed, bool on, uint32_t shift) {
    const int nt = blockDim.x;
    const int chunk = (count + nt - 1) / nt;
    const int k0 = threadIdx.x * chunk;
    if (!on) {
        for (int k = k0; k < min(k0 + chunk, count); k++) g[k] = 0;
        return;
    }
    uint32_t r = fg_jump(seed & 0xffffu, (uint32_t)k0);
    for (int k = k0; k < min(k0 + chunk, count); k++) {
        fg_step(r);
        g[k] = (int16_t)fg_round2((int32_t)c_gauss[(r >> 5) & 2047u], shift);
    }
}

// Causal auto-regressive filter of NT templates of width w, height h in LDS
// (film_grain.rs:255-289, 319-372), run by one wave alone.  Wavefront:
// lane j filters row r0+j and at step t reaches column 3+t-4j, so row y-1
// finished column x+3 one step earlier.  Within one wave LDS accesses are
// performed in program order, so no workgroup barrier is needed between
// steps (a compiler-only fence keeps the accesses in order); templates
// taller than 64 filtered rows take a second pass over the remaining rows,
// which start after every row of the first pass is done.  Each lane slides
// the LAG rows above (columns x-LAG..x+LAG) and the last LAG outputs of its
// own row through registers; per step it reads one new column of the rows
// above.  Coefficient order follows the reference's pos counter: rows
// -LAG..-1 left to right, then row 0 columns -LAG..-1, then (for chroma) the
// luma-average tap.
template <int LAG, int NT, typename LumaTerm>
__device__ void fg_ar(int16_t *const *g, int w, int h, const int32_t (*cf)[25], uint32_t ash, int32_t gmin,
                      int32_t gmax, LumaTerm luma, int wave = 0) {
    constexpr int K = 2 * LAG + 1;
    constexpr int R = LAG > 0 ? LAG : 1;
    if ((int)(threadIdx.x >> 6) != wave) return;
    const int rows_all = h - 3, cols = w - 6;
    const int j = threadIdx.x & 63;
    for (int r0 = 0; r0 < rows_all; r0 += 64) {
        const int rows = min(64, rows_all - r0);
        const int steps = cols + 4 * (rows - 1);
        const int y = 3 + r0 + j;
        int32_t win[NT][R][K];
        int32_t cur[NT][R];
        for (int t = 0; t < steps; t++) {
            const int x = 3 + t - 4 * j;
            if (j < rows && x >= 3 && x < w - 3) {
                if constexpr (LAG > 0) {
                    if (x == 3) {
#pragma unroll
                        for (int n = 0; n < NT; n++) {
#pragma unroll
                            for (int r = 0; r < LAG; r++)
#pragma unroll
                                for (int k = 0; k < K; k++) win[n][r][k] = g[n][(y - LAG + r) * w + (x - LAG + k)];
#pragma unroll
                            for (int k = 0; k < LAG; k++) cur[n][k] = g[n][y * w + (x - LAG + k)];
                        }
                    } else {
#pragma unroll
                        for (int n = 0; n < NT; n++)
#pragma unroll
                            for (int r = 0; r < LAG; r++) {
#pragma unroll
                                for (int k = 0; k < K - 1; k++) win[n][r][k] = win[n][r][k + 1];
                                win[n][r][K - 1] = g[n][(y - LAG + r) * w + (x + LAG)];
                            }
                    }
                }
                const int32_t lt = luma(0, x, y);
#pragma unroll
                for (int n = 0; n < NT; n++) {
                    // four partial sums keep the multiply-add chain short
                    int32_t s4[4] = {0, 0, 0, 0};
                    if constexpr (LAG > 0) {
#pragma unroll
                        for (int r = 0; r < LAG; r++)
#pragma unroll
                            for (int k = 0; k < K; k++) s4[(r * K + k) & 3] += win[n][r][k] * cf[n][r * K + k];
#pragma unroll
                        for (int k = 0; k < LAG; k++) s4[(LAG * K + k) & 3] += cur[n][k] * cf[n][LAG * K + k];
                    }
                    const int32_t s = s4[0] + s4[1] + s4[2] + s4[3] + lt * cf[n][LAG * K + LAG];
                    const int32_t v = clip3(gmin, gmax, (int32_t)g[n][y * w + x] + fg_round2(s, ash));
                    g[n][y * w + x] = (int16_t)v;
                    if constexpr (LAG > 0) {
#pragma unroll
                        for (int k = 0; k < LAG - 1; k++) cur[n][k] = cur[n][k + 1];
                        cur[n][LAG - 1] = v;
                    }
                }
            }
            // next step's reads of the rows above must follow this step's
            // writes: keep the compiler from moving LDS accesses across
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            __builtin_amdgcn_wave_barrier();
        }
    }
}

// One workgroup per frame (blockDim.x >= 128).
__device__ void fg_prep_frame(const FrameJob &J, int16_t *sm /* 4 * 73 * 82 */) {
    G<FgScratch> *S = (G<FgScratch> *)J.fgs;
    const oxav1_film_grain_params &fg = J.fg;
    const int tid = threadIdx.x;
    const uint32_t bd = J.bd, sx = J.sx, sy = J.sy;
    const int32_t gc = 128 << (bd - 8);
    const int32_t gmin = -gc, gmax = (256 << (bd - 8)) - 1 - gc;
    const uint32_t shift = (12u - bd) + fg.grain_scale_shift;
    const int cw = sx ? 44 : 82, ch = sy ? 38 : 73;
    int16_t *L = sm, *Cb = sm + kGrainH * kGrainW, *Cr = Cb + kGrainH * kGrainW;
    const uint32_t np_ = J.np < 3 ? J.np : 3;
    // scaling LUTs (scaling_lookup_init, film_grain.rs:418-466): a thread per
    // entry of each table
    for (int i = tid; i < 3 * 256; i += blockDim.x) {
        const int p = i >> 8, v = i & 255;
        uint8_t val = 0;
        if ((uint32_t)p < np_) {
            if (p == 0 || fg.chroma_scaling_from_luma)
                val = fg_lut_value(fg.num_y_points, fg.point_y_value, fg.point_y_scaling, v);
            else if (p == 1)
                val = fg_lut_value(fg.num_cb_points, fg.point_cb_value, fg.point_cb_scaling, v);
            else
                val = fg_lut_value(fg.num_cr_points, fg.point_cr_value, fg.point_cr_scaling, v);
        }
        S->lut[p][v] = val;
    }
    // per-stripe block offsets (film_grain.rs:626-641): the stripe's LFSR
    // draws split into runs of 16, each run reached by a jump ahead
    {
        const uint32_t nst = S->nstripes, nbl = S->nblocks, nrun = (nbl + 15) / 16;
        for (uint32_t w = tid; w < nst * nrun; w += blockDim.x) {
            const uint32_t s = w / nrun, b0 = (w - s * nrun) * 16;
            uint32_t r = fg.grain_seed;
            r ^= ((s * 37u + 178u) & 255u) << 8;
            r ^= (s * 173u + 105u) & 255u;
            r = fg_jump(r & 0xffffu, b0);
            for (uint32_t b = b0; b < min(b0 + 16, nbl); b++) {
                fg_step(r);
                ((G<uint8_t> *)S->offsets)[s * nbl + b] = (uint8_t)((r >> 8) & 255u);
            }
        }
    }
    if (tid == 0) {
        S->chroma_w = cw;
        S->chroma_h = ch;
    }
    // white noise
    fg_white(L, kGrainH * kGrainW, fg.grain_seed, fg.num_y_points > 0, shift);
    const bool chroma_on_cb = fg.num_cb_points > 0 || fg.chroma_scaling_from_luma;
    const bool chroma_on_cr = fg.num_cr_points > 0 || fg.chroma_scaling_from_luma;
    fg_white(Cb, cw * ch, fg.grain_seed ^ 0xb524u, chroma_on_cb, shift);
    fg_white(Cr, cw * ch, fg.grain_seed ^ 0x49d8u, chroma_on_cr, shift);
    __syncthreads();
    const int lag = fg.ar_coeff_lag;
    const uint32_t ash = fg.ar_coeff_shift;
    // luma AR (film_grain.rs:255-289)
    if (fg.num_y_points > 0) {
        int16_t *gs[1] = {L};
        int32_t cf[1][25];
        for (int i = 0; i < 24; i++) cf[0][i] = (int32_t)fg.ar_coeffs_y_plus_128[i] - 128;
        cf[0][24] = 0;
        auto none = [](int, int, int) { return 0; };
        switch (lag) {
        case 0: fg_ar<0, 1>(gs, kGrainW, kGrainH, cf, ash, gmin, gmax, none); break;
        case 1: fg_ar<1, 1>(gs, kGrainW, kGrainH, cf, ash, gmin, gmax, none); break;
        case 2: fg_ar<2, 1>(gs, kGrainW, kGrainH, cf, ash, gmin, gmax, none); break;
        default: fg_ar<3, 1>(gs, kGrainW, kGrainH, cf, ash, gmin, gmax, none); break;
        }
    }
    __syncthreads();
    // chroma AR (film_grain.rs:319-372): Cb on the first wave and Cr on the
    // second, concurrently; the centre tap adds the co-located luma average
    // (when num_y_points > 0), precomputed for every chroma position
    if (fg.chroma_scaling_from_luma || fg.num_cb_points > 0 || fg.num_cr_points > 0) {
        const bool with_luma = fg.num_y_points > 0;
        int16_t *LT = Cr + kGrainH * kGrainW;  // luma term per chroma position
        const int isx = (int)sx, isy = (int)sy;
        for (int i = tid; i < cw * ch; i += blockDim.x) {
            const int y = i / cw, x = i - y * cw;
            int32_t l = 0;
            if (with_luma && y >= 3 && x >= 3 && x < cw - 3) {
                const int lx = ((x - 3) << isx) + 3, ly = ((y - 3) << isy) + 3;
                for (int a = 0; a <= isy; a++)
                    for (int k = 0; k <= isx; k++) l += (int32_t)L[(ly + a) * kGrainW + (lx + k)];
                l = fg_round2(l, (uint32_t)(isx + isy));
            }
            LT[i] = (int16_t)l;
        }
        __syncthreads();
        int32_t cfb[1][25], cfr[1][25];
        for (int i = 0; i < 25; i++) {
            cfb[0][i] = (int32_t)fg.ar_coeffs_cb_plus_128[i] - 128;
            cfr[0][i] = (int32_t)fg.ar_coeffs_cr_plus_128[i] - 128;
        }
        auto luma = [=](int, int x, int y) -> int32_t { return (int32_t)LT[y * cw + x]; };
        int16_t *gb[1] = {Cb}, *gr[1] = {Cr};
        switch (lag) {
        case 0:
            fg_ar<0, 1>(gb, cw, ch, cfb, ash, gmin, gmax, luma, 0);
            fg_ar<0, 1>(gr, cw, ch, cfr, ash, gmin, gmax, luma, 1);
            break;
        case 1:
            fg_ar<1, 1>(gb, cw, ch, cfb, ash, gmin, gmax, luma, 0);
            fg_ar<1, 1>(gr, cw, ch, cfr, ash, gmin, gmax, luma, 1);
            break;
        case 2:
            fg_ar<2, 1>(gb, cw, ch, cfb, ash, gmin, gmax, luma, 0);
            fg_ar<2, 1>(gr, cw, ch, cfr, ash, gmin, gmax, luma, 1);
            break;
        default:
            fg_ar<3, 1>(gb, cw, ch, cfb, ash, gmin, gmax, luma, 0);
            fg_ar<3, 1>(gr, cw, ch, cfr, ash, gmin, gmax, luma, 1);
            break;
        }
    }
    __syncthreads();
    for (int i = tid; i < kGrainH * kGrainW; i += blockDim.x) S->luma[i] = L[i];
    for (int i = tid; i < cw * ch; i += blockDim.x) {
        S->cb[i] = Cb[i];
        S->cr[i] = Cr[i];
    }
    // the shifted copies, a (plane, shift, row) per kTqW threads
    constexpr int kRowsPerPass = 256 / kTqW;
    const int c = tid % kTqW, rr = tid / kTqW;
    for (int row0 = 0; row0 < 3 * 4 * kGrainH; row0 += kRowsPerPass) {
        const int row = row0 + rr;
        if (rr >= kRowsPerPass || row >= 3 * 4 * kGrainH) continue;
        const int p = row / (4 * kGrainH), k = (row / kGrainH) & 3, r = row % kGrainH;
        const int16_t *t = p == 0 ? L : (p == 1 ? Cb : Cr);
        const int tw = p == 0 ? kGrainW : cw, th = p == 0 ? kGrainH : ch;
        int16_t v = 0;
        if (r < th && c + k < tw) v = t[r * tw + c + k];
        ((G<int16_t> *)S->tq[p][k])[r * kTqW + c] = v;
    }
}

// Grain sample of (stripe s, block b, row i, col j) for a plane.
__device__ __forceinline__ int32_t fg_grain(const G<int16_t> *tmpl, int tw, const G<uint8_t> *offs, uint32_t nblocks,
                                            uint32_t s, uint32_t b, int32_t i, int32_t j, uint32_t psx, uint32_t psy) {
    const uint32_t rnd = offs[s * nblocks + b];
    const int32_t ox = (int32_t)(rnd >> 4), oy = (int32_t)(rnd & 15u);
    const int32_t pox = psx ? 6 + ox : 9 + 2 * ox;
    const int32_t poy = psy ? 6 + oy : 9 + 2 * oy;
    return tmpl[(poy + i) * tw + pox + j];
}

// Value of noiseStripe[s][plane][i][x] after the horizontal overlap blend.
__device__ __forceinline__ int32_t fg_hval(const G<int16_t> *tmpl, int tw, const G<uint8_t> *offs, uint32_t nblocks,
                                           uint32_t s, int32_t i, int32_t x, uint32_t psx, uint32_t psy, bool overlap,
                                           int32_t gmin, int32_t gmax) {
    if (!psx) {
        const uint32_t b = (uint32_t)x >> 5;
        const int32_t j = x & 31;
        int32_t g = fg_grain(tmpl, tw, offs, nblocks, s, b, i, j, psx, psy);
        if (j < 2 && b > 0 && overlap) {
            const int32_t old = fg_grain(tmpl, tw, offs, nblocks, s, b - 1, i, j + 32, psx, psy);
            g = j == 0 ? old * 27 + g * 17 : old * 17 + g * 27;
            g = clip3(gmin, gmax, fg_round2(g, 5));
        }
        return g;
    }
    const uint32_t b = (uint32_t)x >> 4;
    const int32_t j = x & 15;
    int32_t g = fg_grain(tmpl, tw, offs, nblocks, s, b, i, j, psx, psy);
    if (j == 0 && b > 0 && overlap) {
        const int32_t old = fg_grain(tmpl, tw, offs, nblocks, s, b - 1, i, 16, psx, psy);
        g = clip3(gmin, gmax, fg_round2(old * 23 + g * 22, 5));
    }
    return g;
}

// noiseImage[plane][y][x] (film_grain.rs:705-739)
__device__ __forceinline__ int32_t fg_noise(const G<int16_t> *tmpl, int tw, const G<uint8_t> *offs, uint32_t nblocks,
                                            int32_t y, int32_t x, uint32_t psx, uint32_t psy, bool overlap,
                                            int32_t gmin, int32_t gmax) {
    const uint32_t s = (uint32_t)y >> (5 - psy);
    const int32_t i = y - (int32_t)(s << (5 - psy));
    int32_t g = fg_hval(tmpl, tw, offs, nblocks, s, i, x, psx, psy, overlap, gmin, gmax);
    if (!psy) {
        if (i < 2 && s > 0 && overlap) {
            const int32_t old = fg_hval(tmpl, tw, offs, nblocks, s - 1, i + 32, x, psx, psy, overlap, gmin, gmax);
            g = i == 0 ? old * 27 + g * 17 : old * 17 + g * 27;
            g = clip3(gmin, gmax, fg_round2(g, 5));
        }
    } else if (i < 1 && s > 0 && overlap) {
        const int32_t old = fg_hval(tmpl, tw, offs, nblocks, s - 1, i + 16, x, psx, psy, overlap, gmin, gmax);
        g = clip3(gmin, gmax, fg_round2(old * 23 + g * 22, 5));
    }
    return g;
}

// scale_lut (film_grain.rs:543-555)
__device__ __forceinline__ int32_t fg_scale(const uint8_t *t, int32_t index, uint32_t bd) {
    const uint32_t shift = bd - 8;
    const uint32_t x = (uint32_t)(index >> shift);
    if (bd == 8 || x >= 255) return t[min(x, 255u)];
    const int32_t rem = index - (int32_t)(x << shift);
    const int32_t a = t[x], b = t[x + 1];
    return a + fg_round2((b - a) * rem, shift);
}

// Four consecutive samples of T as one vector (raw, packed as loaded).
typedef uint32_t fg_u32x2 __attribute__((ext_vector_type(2)));
typedef int32_t fg_i32x4 __attribute__((ext_vector_type(4)));
template <typename T>
struct FgV;
template <>
struct FgV<uint8_t> {
    using type = uint32_t;
};
template <>
struct FgV<uint16_t> {
    using type = fg_u32x2;
};
template <>
struct FgV<int32_t> {
    using type = fg_i32x4;
};
// sample k of a raw vector
template <typename T>
__device__ __forceinline__ int32_t fg_at(const typename FgV<T>::type &v, int k) {
    if constexpr (sizeof(T) == 1)
        return (int32_t)((v >> (8 * k)) & 255u);
    else if constexpr (sizeof(T) == 2)
        return (int32_t)(((k < 2 ? v.x : v.y) >> (16 * (k & 1))) & 0xffffu);
    else
        return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w));
}
// four samples packed into a raw vector
template <typename T>
__device__ __forceinline__ typename FgV<T>::type fg_pack(const int32_t (&t)[4]) {
    if constexpr (sizeof(T) == 1)
        return (uint32_t)t[0] | (uint32_t)t[1] << 8 | (uint32_t)t[2] << 16 | (uint32_t)t[3] << 24;
    else if constexpr (sizeof(T) == 2)
        return fg_u32x2{(uint32_t)t[0] | (uint32_t)t[1] << 16, (uint32_t)t[2] | (uint32_t)t[3] << 16};
    else
        return fg_i32x4{t[0], t[1], t[2], t[3]};
}
// Four samples of a plane row at p: one vector load where vec (p aligned to
// four samples, all four columns inside), else the first n one by one.
template <typename T>
__device__ __forceinline__ typename FgV<T>::type fg_ld4(const G<T> *p, bool vec, int n) {
    if (vec) return *(const G<typename FgV<T>::type> *)p;
    int32_t t[4];
#pragma unroll
    for (int k = 0; k < 4; k++) t[k] = k < n ? (int32_t)p[k] : 0;
    return fg_pack<T>(t);
}
template <typename T>
__device__ __forceinline__ void fg_st4(G<T> *p, bool vec, int n, const int32_t (&v)[4]) {
    if (vec) {
        *(G<typename FgV<T>::type> *)p = fg_pack<T>(v);
    } else {
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (k < n) p[k] = (T)v[k];
    }
}
// four consecutive template values at (row r, column c) of plane p's
// template, from the shifted copy c & 3
__device__ __forceinline__ fg_u32x2 fg_tq4(const G<int16_t> *tq, int32_t r, int32_t c) {
    return *(const G<fg_u32x2> *)(tq + ((uint32_t)(c & 3) * (kGrainH * kTqW) + (uint32_t)(r * kTqW + (c & ~3))));
}
__device__ __forceinline__ int32_t fg_tq_at(const fg_u32x2 &v, int k) {
    const uint32_t w = k < 2 ? v.x : v.y;
    return (k & 1) ? (int32_t)w >> 16 : (int32_t)(int16_t)(w & 0xffffu);
}

// Clamp of v to [lo, hi] (lo <= hi) as one v_med3_i32.
__device__ __forceinline__ int32_t fg_med3(int32_t v, int32_t lo, int32_t hi) {
    int32_t r;
    asm("v_med3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(v), "v"(lo), "v"(hi));
    return r;
}
// a * b: u8/u16 frames keep every product of the blend within the 24-bit
// multiplier's operand range (samples < 2^12, noise and weights < 2^12);
// the int32 drop-in path keeps the reference's 32-bit products.
template <typename T>
__device__ __forceinline__ int32_t fg_mul(int32_t a, int32_t b) {
    if constexpr (sizeof(T) == 4)
        return a * b;
    else
        return __mul24(a, b);
}
// scale_lut (film_grain.rs:543-555) without a branch: the top entry (or any
// index whose bucket is 255 or more) has its next entry equal to itself, so
// the interpolation adds nothing there.  rnd = Round2's constant of shift.
template <typename T>
__device__ __forceinline__ int32_t fg_scale_at(const uint8_t *t, int32_t index, uint32_t shift, int32_t rnd) {
    const uint32_t x = min((uint32_t)(index >> shift), 255u);
    const int32_t a = t[x], b = t[min(x + 1, 255u)];
    const int32_t rem = index - (int32_t)(x << shift);
    return a + ((fg_mul<T>(b - a, rem) + rnd) >> shift);
}

// Blend one tile (rows [y0, y1), cols [x0, x1), at most 64 x 64, tile
// corners on multiples of 64) of one plane: in (pre-grain) -> out
// (add_noise_synthesis, film_grain.rs:570-829).  A lane owns four
// consecutive columns of one row (16 lanes per row); each wave walks a band
// of 16 rows, four rows at a time, with every load of the band issued
// before the first store.  Bands never straddle a noise stripe (32 or 16
// rows) and a lane's columns never straddle a noise block (32 or 16
// columns), so the block offsets are resolved once per lane, the horizontal
// overlap blend only touches the lanes at a block's first columns and the
// vertical one the first rows of a stripe.
template <typename T, class JT = const FrameJob>
__device__ __forceinline__ void fg_blend_tile(JT &J, uint32_t plane, int32_t y0, int32_t y1, int32_t x0, int32_t x1,
                              const uint8_t *lut /* LDS [3][256] */) {
    const G<FgScratch> *S = (const G<FgScratch> *)J.fgs;
    auto &fg = J.fg;
    const uint32_t bd = J.bd;
    const uint32_t psx = plane ? J.sx : 0, psy = plane ? J.sy : 0;
    const int32_t gc = 128 << (bd - 8);
    const int32_t gmin = -gc, gmax = (256 << (bd - 8)) - 1 - gc;
    int32_t minv, maxl, maxc;
    if (fg.clip_to_restricted_range) {
        minv = 16 << (bd - 8);
        maxl = 235 << (bd - 8);
        maxc = J.mc == 0 ? maxl : (240 << (bd - 8));
    } else {
        minv = 0;
        maxl = maxc = (256 << (bd - 8)) - 1;
    }
    const uint32_t gs = fg.grain_scaling;
    const int32_t rndg = gs ? 1 << (gs - 1) : 0;              // Round2 constants
    const uint32_t shift = bd - 8;
    const int32_t rnds = shift ? 1 << (shift - 1) : 0;
    const bool overlap = fg.overlap_flag != 0;
    const G<T> *in = plane_ptr<T>(J.ref_out[plane]);
    G<T> *out = plane_ptr<T>(J.out[plane]);
    const uint32_t ip = J.ref_pitch[plane], op = J.out_pitch[plane];
    const int32_t w = (int32_t)J.out_cols[0];
    // the wave index is uniform in a wave: a scalar, so the band's geometry is too
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int rs = lane >> 4;                         // row within a chunk of four
    const int32_t x = x0 + 4 * (lane & 15);
    const int32_t yb0 = y0 + 16 * wv, yb1 = min(y1, yb0 + 16);
    if (yb0 >= yb1) return;
    const int n = min(4, x1 - x);                     // this lane's columns (<= 0: none)
    const bool vin = (ip & 3u) == 0 && ((uintptr_t)in & (4 * sizeof(T) - 1)) == 0 && n == 4;
    const bool vout = (op & 3u) == 0 && ((uintptr_t)out & (4 * sizeof(T) - 1)) == 0 && n == 4;
    // A job without grain (apply_grain == 0, out != ref_out) has no grain
    // scratch: it only copies.
    bool active;
    if (!J.apply_fg) {
        active = false;
    } else if (plane == 0) {
        active = fg.num_y_points > 0;
    } else {
        const bool any = fg.num_cb_points > 0 || fg.num_cr_points > 0 || fg.chroma_scaling_from_luma;
        active = J.np >= 3 && any &&
                 (plane == 1 ? (fg.num_cb_points > 0 || fg.chroma_scaling_from_luma)
                             : (fg.num_cr_points > 0 || fg.chroma_scaling_from_luma));
    }
    using V = typename FgV<T>::type;
    if (!active) {
        V v[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int32_t y = yb0 + 4 * c + rs;
            if (n > 0 && y < yb1) v[c] = fg_ld4<T>(in + ((uint32_t)y * ip + (uint32_t)x), vin, n);
        }
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int32_t y = yb0 + 4 * c + rs;
            if (n > 0 && y < yb1) {
                int32_t t[4];
#pragma unroll
                for (int k = 0; k < 4; k++) t[k] = fg_at<T>(v[c], k);
                fg_st4<T>(out + ((uint32_t)y * op + (uint32_t)x), vout, n, t);
            }
        }
        return;
    }
    const G<int16_t> *tmpl = plane == 0 ? S->luma : (plane == 1 ? S->cb : S->cr);
    const G<int16_t> *tq = S->tq[plane][0];
    const int tw = plane == 0 ? kGrainW : (int)S->chroma_w;
    // noise block / stripe geometry (film_grain.rs:626-651, 705-739)
    const uint32_t bsh = 5 - psx, ssh = 5 - psy;       // log2 block width / stripe height
    const uint32_t b = (uint32_t)max(x, 0) >> bsh;
    const int32_t j = x & ((1 << bsh) - 1);            // a multiple of four
    const uint32_t st = (uint32_t)yb0 >> ssh;
    const int32_t ib = yb0 - (int32_t)(st << ssh);       // row of yb0 inside its stripe (uniform)
    const G<uint8_t> *offs = (const G<uint8_t> *)S->offsets;
    const uint32_t nblocks = S->nblocks;
    // (column, row) of a block's noise in the template
    auto block_pos = [&](uint32_t ss, uint32_t bb, int32_t &pox, int32_t &poy) {
        const uint32_t rnd = offs[ss * nblocks + bb];
        const int32_t ox = (int32_t)(rnd >> 4), oy = (int32_t)(rnd & 15u);
        pox = psx ? 6 + ox : 9 + 2 * ox;
        poy = psy ? 6 + oy : 9 + 2 * oy;
    };
    // horizontal overlap: the first 2 (1 when subsampled) columns of a block
    // blend with the previous block's columns 32, 33 (16)
    const int nh = psx ? 1 : 2;
    const bool hov = overlap && b > 0 && j == 0;
    int32_t pcx = 0, pcy = 0, plx = 0, ply = 0;
    if (n > 0) block_pos(st, b, pcx, pcy);
    if (n > 0 && hov) block_pos(st, b - 1, plx, ply);
    plx += 1 << bsh;
    // vertical overlap: the first 2 (1) rows of a stripe blend with rows 32,
    // 33 (16) of the previous stripe's noise; a band holds them only when it
    // starts the stripe (ib == 0)
    const int32_t vrows = (overlap && st > 0 && ib == 0) ? (psy ? 1 : 2) : 0;
    int32_t pux = 0, puy = 0, pulx = 0, puly = 0;
    if (vrows > 0 && n > 0) {
        block_pos(st - 1, b, pux, puy);
        if (hov) block_pos(st - 1, b - 1, pulx, puly);
        puy += 1 << ssh;
        puly += 1 << ssh;
        pulx += 1 << bsh;
    }
    const uint8_t *lt = lut + 256 * plane;
    const G<T> *ly = plane_ptr<T>(J.ref_out[0]);
    const uint32_t lp = J.ref_pitch[0];
    const bool from_luma = fg.chroma_scaling_from_luma != 0;
    const int32_t lm = plane == 1 ? fg.cb_luma_mult : fg.cr_luma_mult;
    const int32_t mm = plane == 1 ? fg.cb_mult : fg.cr_mult;
    const int32_t coff = ((plane == 1 ? fg.cb_offset : fg.cr_offset) - 256) * (1 << (bd - 8));
    const int32_t maxs = (1 << bd) - 1;
    // co-located luma of the lane's four columns: columns x << psx .. ,
    // the odd one of each pair clamped to the luma width
    const int32_t lx = x << psx;
    const bool vl = (lp & 3u) == 0 && ((uintptr_t)ly & (4 * sizeof(T) - 1)) == 0 && lx + (4 << psx) <= w;
    // The band is blended a chunk of four rows at a time: every load of a
    // chunk is issued before its first store and kept as loaded.
    const int32_t hwo0 = psx ? 23 : 27, hwn0 = psx ? 22 : 17;  // column 0 of a block (row 0 of a stripe)
    fg_u32x2 vor = {0u, 0u};  // the previous stripe's overlap row (first chunk, rows < vrows)
    int32_t vl2[2] = {0, 0};
    if (vrows > 0 && n > 0 && rs < vrows) {
        vor = fg_tq4(tq, puy + rs, pux + j);
        if (hov)
            for (int k = 0; k < nh; k++) vl2[k] = tmpl[(puly + rs) * tw + pulx + k];
    }
    constexpr int kCh = 1;  // chunks per step
#pragma unroll 1
    for (int h = 0; h < 4; h += kCh) {
        fg_u32x2 nzr[kCh];
        V org[kCh], lar[kCh], lbr[kCh];
        int32_t hl[kCh][2] = {};
#pragma unroll
        for (int cc = 0; cc < kCh; cc++) {
            const int c = h + cc;
            const int32_t y = yb0 + 4 * c + rs;
            const bool ok = n > 0 && y < yb1;
            const int32_t i = ib + 4 * c + rs;
            if (ok) {
                nzr[cc] = fg_tq4(tq, pcy + i, pcx + j);
                org[cc] = fg_ld4<T>(in + ((uint32_t)y * ip + (uint32_t)x), vin, n);
                if (hov)
                    for (int k = 0; k < nh; k++) hl[cc][k] = tmpl[(ply + i) * tw + plx + k];
            }
            if (plane && ok) {
                const G<T> *lrow = ly + (uint32_t)(y << psy) * lp;
                if (vl) {
                    lar[cc] = fg_ld4<T>(lrow + lx, true, 4);
                    if (psx) lbr[cc] = fg_ld4<T>(lrow + lx + 4, true, 4);
                } else {
                    // at the right edge: luma columns clamped to the luma width
                    int32_t t0[4], t1[4];
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        t0[k] = (int32_t)lrow[min(lx + k, w - 1)];
                        t1[k] = psx ? (int32_t)lrow[min(lx + 4 + k, w - 1)] : 0;
                    }
                    lar[cc] = fg_pack<T>(t0);
                    lbr[cc] = fg_pack<T>(t1);
                }
            }
        }
#pragma unroll
        for (int cc = 0; cc < kCh; cc++) {
            const int c = h + cc;
            const int32_t y = yb0 + 4 * c + rs;
            if (n <= 0 || y >= yb1) continue;
            const int32_t i = ib + 4 * c + rs;
            // the chunk's noise with the overlap blends: (27, 17) at column /
            // row 0, (17, 27) at 1; (23, 22) when subsampled
            int32_t v[4];
#pragma unroll
            for (int k = 0; k < 4; k++) v[k] = fg_tq_at(nzr[cc], k);
            if (hov) {
#pragma unroll
                for (int k = 0; k < 2; k++)
                    if (k < nh)
                        v[k] = fg_med3((__mul24(hl[cc][k], k == 0 ? hwo0 : 17) + __mul24(v[k], k == 0 ? hwn0 : 27) + 16) >> 5,
                                       gmin, gmax);
            }
            if (c == 0 && rs < vrows) {  // the stripe's first rows: i == rs
                int32_t o[4];
#pragma unroll
                for (int k = 0; k < 4; k++) o[k] = fg_tq_at(vor, k);
                if (hov) {
#pragma unroll
                    for (int k = 0; k < 2; k++)
                        if (k < nh)
                            o[k] = fg_med3((__mul24(vl2[k], k == 0 ? hwo0 : 17) + __mul24(o[k], k == 0 ? hwn0 : 27) + 16) >> 5,
                                           gmin, gmax);
                }
                const int32_t wo = psy ? 23 : (rs == 0 ? 27 : 17), wn = psy ? 22 : (rs == 0 ? 17 : 27);
#pragma unroll
                for (int k = 0; k < 4; k++) v[k] = fg_med3((__mul24(o[k], wo) + __mul24(v[k], wn) + 16) >> 5, gmin, gmax);
            }
            int32_t res[4];
            if (plane == 0) {
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int32_t orig = fg_at<T>(org[cc], k);
                    const int32_t sc = fg_scale_at<T>(lt, orig, shift, rnds);
                    res[k] = fg_med3(orig + ((fg_mul<T>(sc, v[k]) + rndg) >> gs), minv, maxl);
                }
            } else {
                // chroma: scaled by the co-located pre-noise luma
                // (film_grain.rs:757-808): luma columns lx + 2k, lx + 2k + 1
                // (lx + k without subsampling)
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int32_t orig = fg_at<T>(org[cc], k);
                    int32_t avg;
                    if (psx) {
                        const V &lv = k < 2 ? lar[cc] : lbr[cc];
                        avg = (fg_at<T>(lv, 2 * (k & 1)) + fg_at<T>(lv, 2 * (k & 1) + 1) + 1) >> 1;
                    } else {
                        avg = fg_at<T>(lar[cc], k);
                    }
                    int32_t merged;
                    if (from_luma) {
                        merged = avg;
                    } else {
                        const int32_t comb = fg_mul<T>(avg, lm - 128) + fg_mul<T>(orig, mm - 128);
                        merged = fg_med3((comb >> 6) + coff, 0, maxs);
                    }
                    const int32_t sc = fg_scale_at<T>(lt, merged, shift, rnds);
                    res[k] = fg_med3(orig + ((fg_mul<T>(sc, v[k]) + rndg) >> gs), minv, maxc);
                }
            }
            fg_st4<T>(out + ((uint32_t)y * op + (uint32_t)x), vout, n, res);
        }
    }
}

}  // namespace oxav1
