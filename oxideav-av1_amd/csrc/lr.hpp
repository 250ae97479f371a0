// §7.17 loop restoration on gfx950: one workgroup per (plane, 64-luma-row
// stripe, column tile), Wiener or self-guided chosen per tile.
//
// Reference: loop_restoration.rs:268-992.  The reference filters 4x4 blocks
// one at a time and fetches every tap through get_source_sample
// (loop_restoration.rs:862-883), which routes rows above/below the block's
// stripe to the pre-CDEF frame (+-2 rows) and everything else to the CDEF
// frame.  Inside one stripe that fetch is a pure function of (x, y), so the
// workgroup materialises the routed window [stripe-3, stripe+3] x
// [tile-3, tile+3] in LDS once and every block of the stripe reads it.
// Restoration units never straddle a stripe (unit rows start at
// unitSize*k - 8 luma rows, stripes at 64k - 8) and tile widths divide the
// unit size, so one tile has one unit's parameters.  Self-guided A/B maps
// are built per stripe over [stripe-1, stripe+1] with the stripe's routing
// (the reference computes each block's A/B halo with that block's routing,
// which is the same map); pass 0 only needs odd rows (loop_restoration.rs
// :686-699 weights even rows 0, and block rows start even).
#pragma once

#include "common.hpp"

namespace oxav1 {

constexpr int kLrMaxW = 64;                 // max tile width (plane samples)
constexpr int kLrMaxH = 64;                 // max stripe height
constexpr int kLrSw = kLrMaxW + 6;          // source window width
constexpr int kLrSh = kLrMaxH + 6;          // source window height
constexpr int kLrAw = kLrMaxW + 2;          // A/B width
constexpr int kLrAh = kLrMaxH + 6;          // A/B height (pass 1); Wiener rows
constexpr int kLrAh0 = kLrMaxH / 2 + 1;     // odd rows only (pass 0)

__constant__ int32_t c_sgr_params[16][4];   // filled from kSgrParams at init
__constant__ uint16_t c_sgr_a2[256];        // a2 for 0 < z < 255 (loop_restoration.rs:648-654)

template <typename L>
struct LrSmem {
    L src[kLrSh * kLrSw];
    union {
        int32_t inter[kLrAh * kLrAw];  // Wiener horizontal pass, (H+6) rows
        struct {
            uint16_t a1[(kLrMaxH + 2) * kLrAw];
            int32_t b1[(kLrMaxH + 2) * kLrAw];
            uint16_t a0[kLrAh0 * kLrAw];
            int32_t b0[kLrAh0 * kLrAw];
        } sg;
    } u;
};

struct LrTileGeom {
    int32_t ss, se;      // stripe start / end rows (may be < 0 / > plane end)
    int32_t pex, pey;    // PlaneEndX / PlaneEndY
    int32_t ys, ye;      // output rows (inclusive)
    int32_t xs, xe;      // output cols (inclusive)
    int32_t wy0, wx0;    // window origin (virtual coordinates)
    int32_t ww, wh;      // window size
};

// get_source_sample (loop_restoration.rs:862-902) for a virtual (x, y).
template <typename T>
__device__ __forceinline__ int32_t lr_fetch(const LrTileGeom &g, const G<T> *__restrict__ cur, uint32_t cpitch,
                                            int32_t crows, int32_t ccols, const G<T> *__restrict__ cdf,
                                            uint32_t dpitch, int32_t drows, int32_t dcols, int32_t x, int32_t y) {
    x = clip3(0, max(g.pex, 0), x);
    y = clip3(0, max(g.pey, 0), y);
    const G<T> *b;
    uint32_t pitch;
    int32_t rows, cols;
    if (y < g.ss) {
        y = max(y, g.ss - 2);
        b = cur, pitch = cpitch, rows = crows, cols = ccols;
    } else if (y > g.se) {
        y = min(y, g.se + 2);
        b = cur, pitch = cpitch, rows = crows, cols = ccols;
    } else {
        b = cdf, pitch = dpitch, rows = drows, cols = dcols;
    }
    if (rows == 0 || cols == 0) return 0;
    y = clip3(0, rows - 1, y);
    x = clip3(0, cols - 1, x);
    return (int32_t)b[(size_t)y * pitch + x];
}

// Self-guided A/B at one position (loop_restoration.rs:615-665) from the
// LDS window; (wy, wx) is the window coordinate of the centre.
template <typename L>
__device__ __forceinline__ void lr_ab(const L *src, int32_t wy, int32_t wx, int32_t r, int32_t n, uint32_t s,
                                      int32_t one_over_n, uint32_t bd, uint16_t *A, int32_t *B) {
    int32_t a = 0, b = 0;
    for (int32_t dy = -r; dy <= r; dy++) {
        const L *row = src + (wy + dy) * kLrSw + wx;
        for (int32_t dx = -r; dx <= r; dx++) {
            const int32_t c = (int32_t)row[dx];
            a += c * c;
            b += c;
        }
    }
    const uint32_t sha = 2 * (bd - 8), shb = bd - 8;
    const int32_t ar = sha ? (int32_t)(((uint32_t)a + (1u << (sha - 1))) >> sha) : a;
    const int32_t dr = shb ? ((b + (1 << (shb - 1))) >> shb) : b;
    int32_t p = ar * n - dr * dr;
    p = p < 0 ? 0 : p;
    const uint32_t z = (uint32_t)(((uint64_t)(uint32_t)p * s + (1u << 19)) >> 20);
    const int32_t a2 = z >= 255 ? 256 : (z == 0 ? 1 : (int32_t)c_sgr_a2[z]);
    // b2 = (256 - a2) * b * oneOverN is non-negative and < 2^32 at every depth
    const uint32_t b2 = (uint32_t)(256 - a2) * (uint32_t)b * (uint32_t)one_over_n;
    *A = (uint16_t)a2;
    *B = (int32_t)((b2 + 2048u) >> 12);
}

// One tile: (plane, stripe, column tile) -> out (cropped frame coordinates
// equal plane coordinates).  Writes rows [ys, ye] x cols [xs, xe].
template <typename T>
__device__ void lr_tile(const FrameJob &J, uint32_t plane, uint32_t stripe, uint32_t tile,
                        LrSmem<typename LdsOf<T>::type> &S, G<T> *__restrict__ out, uint32_t opitch) {
    using L = typename LdsOf<T>::type;
    const uint32_t sx = plane ? J.sx : 0, sy = plane ? J.sy : 0;
    LrTileGeom g;
    g.ss = (-8 + (int32_t)stripe * 64) >> sy;
    g.se = g.ss + (64 >> sy) - 1;
    g.pex = (int32_t)((J.uw + sx) >> sx) - 1;
    g.pey = (int32_t)((J.fh + sy) >> sy) - 1;
    g.ys = max(g.ss, 0);
    g.ye = min(g.se, g.pey);
    const int32_t tw = (int32_t)J.lr_tile_w[plane];
    g.xs = (int32_t)tile * tw;
    g.xe = min(g.xs + tw - 1, g.pex);
    if (g.ys > g.ye || g.xs > g.xe) return;
    const G<T> *cur = plane_ptr<T>(J.up_deb[plane]);
    const G<T> *cdf = plane_ptr<T>(J.up_cdef[plane]);
    const uint32_t cpitch = J.up_deb_pitch[plane], dpitch = J.up_cdef_pitch[plane];
    const int32_t rows = (int32_t)J.lr_rows[plane], cols = (int32_t)J.lr_cols[plane];
    const int32_t crows = (int32_t)J.lrc_rows[plane], ccols = (int32_t)J.lrc_cols[plane];
    const int tid = threadIdx.x, nt = blockDim.x;
    // writes are dropped outside the output buffer (loop_restoration.rs:814-818)
    g.ye = min(g.ye, (int32_t)J.out_rows[plane] - 1);
    g.xe = min(g.xe, (int32_t)J.out_cols[plane] - 1);
    if (g.ys > g.ye || g.xs > g.xe) return;
    const int32_t H = g.ye - g.ys + 1, W = g.xe - g.xs + 1;

    // unit of this tile (loop_restoration.rs:395-405)
    uint32_t type = J.uses_lr ? J.lr_type[plane] : OXAV1_RESTORE_NONE;
    oxav1_lr_unit u{};
    if (type != OXAV1_RESTORE_NONE) {
        const uint32_t us = J.lr_unit_size[plane];
        const uint32_t ur = min(J.lr_unit_rows[plane] - 1, (uint32_t)(g.ys + (8 >> sy)) / us);
        const uint32_t uc = min(J.lr_unit_cols[plane] - 1, (uint32_t)g.xs / us);
        u = J.lr_units[plane] ? J.lr_units[plane][(size_t)ur * J.lr_unit_cols[plane] + uc] : oxav1_lr_unit{};
        type = u.type;
    }
    if (type != OXAV1_RESTORE_WIENER && type != OXAV1_RESTORE_SGRPROJ) {
        if (!J.lr_copy_none) return;
        for (int i = tid; i < H * W; i += nt) {
            const int yy = i / W, xx = i - yy * W;
            const int32_t y = g.ys + yy, x = g.xs + xx;
            out[(size_t)y * opitch + x] = cdf[(size_t)y * dpitch + x];
        }
        return;
    }
    // routed source window
    g.wy0 = g.ys - 3;
    g.wx0 = g.xs - 3;
    g.wh = H + 6;
    g.ww = W + 6;
    for (int i = tid; i < g.wh * g.ww; i += nt) {
        const int yy = i / g.ww, xx = i - yy * g.ww;
        S.src[yy * kLrSw + xx] =
            (L)lr_fetch<T>(g, cur, cpitch, crows, ccols, cdf, dpitch, rows, cols, g.wx0 + xx, g.wy0 + yy);
    }
    __syncthreads();
    const int32_t bd = J.bd;
    const int32_t maxv = (1 << bd) - 1;
    if (type == OXAV1_RESTORE_WIENER) {
        // wiener_filter (loop_restoration.rs:742-821)
        int32_t vf[7], hf[7];
        vf[3] = hf[3] = 128;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            vf[i] = vf[6 - i] = u.wiener[0][i];
            hf[i] = hf[6 - i] = u.wiener[1][i];
            vf[3] -= 2 * u.wiener[0][i];
            hf[3] -= 2 * u.wiener[1][i];
        }
        const uint32_t r0 = bd == 12 ? 5 : 3, r1 = bd == 12 ? 9 : 11;
        const int32_t offset = 1 << (bd + 7 - (int32_t)r0 - 1);
        const int32_t limit = (1 << (bd + 1 + 7 - (int32_t)r0)) - 1;
        int32_t *inter = S.u.inter;  // (H+6) x W, stride kLrAw
        for (int i = tid; i < (H + 6) * W; i += nt) {
            const int rr = i / W, c = i - rr * W;
            const L *s = S.src + rr * kLrSw + c;  // window col c == virtual xs + c - 3
            int32_t acc = 0;
#pragma unroll
            for (int t = 0; t < 7; t++) acc += hf[t] * (int32_t)s[t];
            inter[rr * kLrAw + c] = clip3(-offset, limit - offset, (acc + (1 << (r0 - 1))) >> r0);
        }
        __syncthreads();
        for (int i = tid; i < H * W; i += nt) {
            const int rr = i / W, c = i - rr * W;
            int32_t acc = 0;
#pragma unroll
            for (int t = 0; t < 7; t++) acc += vf[t] * inter[(rr + t) * kLrAw + c];
            const int32_t v = (acc + (1 << (r1 - 1))) >> r1;
            out[(size_t)(g.ys + rr) * opitch + (g.xs + c)] = (T)clip3(0, maxv, v);
        }
        return;
    }
    // self_guided_filter (loop_restoration.rs:485-579) + box_filter
    const uint32_t set = u.sgr_set < 15 ? u.sgr_set : 15;
    const int32_t r0 = c_sgr_params[set][0], e0 = c_sgr_params[set][1];
    const int32_t r1 = c_sgr_params[set][2], e1 = c_sgr_params[set][3];
    const int32_t w0 = u.sgr_xqd[0], w1 = u.sgr_xqd[1];
    const int32_t w2 = 128 - w0 - w1;
    auto params = [](int32_t r, int32_t e, int32_t *n, uint32_t *s, int32_t *oon) {
        *n = (2 * r + 1) * (2 * r + 1);
        const int64_t n2e = (int64_t)(*n) * (*n) * e;
        *s = (uint32_t)((((int64_t)1 << 20) + n2e / 2) / (n2e > 1 ? n2e : 1));
        *oon = ((1 << 12) + *n / 2) / *n;
    };
    // A/B rows: virtual ys-1 .. ye+1 -> index 0 .. H+1; window row = index + 2
    const int AW = W + 2;
    if (r0 != 0) {
        int32_t n;
        uint32_t s;
        int32_t oon;
        params(r0, e0, &n, &s, &oon);
        // odd absolute rows only: virtual row vy odd; index k = (vy - (ys-1)) >> 1
        // ys is even, so ys-1 is odd and vy = ys-1+2k.
        const int AH0 = (H + 2 + 1) / 2;
        for (int i = tid; i < AH0 * AW; i += nt) {
            const int k = i / AW, c = i - k * AW;
            const int wy = 2 * k + 2, wx = c + 2;
            lr_ab<L>(S.src, wy, wx, r0, n, s, oon, (uint32_t)bd, &S.u.sg.a0[k * kLrAw + c], &S.u.sg.b0[k * kLrAw + c]);
        }
    }
    if (r1 != 0) {
        int32_t n;
        uint32_t s;
        int32_t oon;
        params(r1, e1, &n, &s, &oon);
        for (int i = tid; i < (H + 2) * AW; i += nt) {
            const int rr = i / AW, c = i - rr * AW;
            lr_ab<L>(S.src, rr + 2, c + 2, r1, n, s, oon, (uint32_t)bd, &S.u.sg.a1[rr * kLrAw + c], &S.u.sg.b1[rr * kLrAw + c]);
        }
    }
    __syncthreads();
    for (int i = tid; i < H * W; i += nt) {
        const int rr = i / W, c = i - rr * W;
        const int32_t uval = (int32_t)S.src[(rr + 3) * kLrSw + (c + 3)];
        const int32_t u4 = uval << 4;
        int32_t f0 = u4, f1 = u4;
        if (r0 != 0) {
            // row parity relative to the block == absolute parity (block rows are even)
            const int32_t vy = g.ys + rr;
            int32_t as, bs;
            uint32_t sh;
            if (vy & 1) {
                const int k = (vy - (g.ys - 1)) >> 1;  // this (odd) row
                const uint16_t *a = S.u.sg.a0 + k * kLrAw + c;
                const int32_t *b = S.u.sg.b0 + k * kLrAw + c;
                as = 5 * a[0] + 6 * a[1] + 5 * a[2];
                bs = 5 * b[0] + 6 * b[1] + 5 * b[2];
                sh = 4;
            } else {
                const int k = (vy - 1 - (g.ys - 1)) >> 1;  // rows vy-1 and vy+1
                const uint16_t *a = S.u.sg.a0 + k * kLrAw + c;
                const int32_t *b = S.u.sg.b0 + k * kLrAw + c;
                const uint16_t *a2 = a + kLrAw;
                const int32_t *b2 = b + kLrAw;
                as = 5 * (a[0] + a[2] + a2[0] + a2[2]) + 6 * (a[1] + a2[1]);
                bs = 5 * (b[0] + b[2] + b2[0] + b2[2]) + 6 * (b[1] + b2[1]);
                sh = 5;
            }
            const int32_t v = as * uval + bs;
            const uint32_t ts = 8 + sh - 4;
            f0 = (v + (1 << (ts - 1))) >> ts;
        }
        if (r1 != 0) {
            const uint16_t *a = S.u.sg.a1 + rr * kLrAw + c;
            const int32_t *b = S.u.sg.b1 + rr * kLrAw + c;
            const int32_t as = 3 * (a[0] + a[2] + a[2 * kLrAw] + a[2 * kLrAw + 2]) +
                               4 * (a[1] + a[kLrAw] + a[kLrAw + 1] + a[kLrAw + 2] + a[2 * kLrAw + 1]);
            const int32_t bs = 3 * (b[0] + b[2] + b[2 * kLrAw] + b[2 * kLrAw + 2]) +
                               4 * (b[1] + b[kLrAw] + b[kLrAw + 1] + b[kLrAw + 2] + b[2 * kLrAw + 1]);
            const int32_t v = as * uval + bs;
            f1 = (v + (1 << 8)) >> 9;
        }
        int32_t v = w1 * u4;
        v += r0 != 0 ? w0 * f0 : w0 * u4;
        v += r1 != 0 ? w2 * f1 : w2 * u4;
        const int32_t sres = (v + (1 << 10)) >> 11;
        out[(size_t)(g.ys + rr) * opitch + (g.xs + c)] = (T)clip3(0, maxv, sres);
    }
}

}  // namespace oxav1
