// §7.14 deblocking on gfx950: one workgroup per 64x64 plane tile, both
// passes fused, out of place (src -> deb).
//
// Reference: loop_filter.rs:271-904.  The reference walks every edge of a
// plane sequentially and in place: all vertical edges (pass 0), then all
// horizontal edges (pass 1).  For legal block/transform geometry the used
// read/write footprints of different edges of one pass never overlap
// (an edge of filter size s reads +-s/2 and writes +-(s/2-1) inside the two
// transforms it separates), so every edge of a pass can be filtered at
// once.  A tile therefore loads its 64x64 region plus a 12-sample halo into
// LDS, runs pass 0 on every loaded row for the vertical edges whose writes
// reach the tile's columns, runs pass 1 for the horizontal edges whose writes
// reach the tile's rows, and stores only the tile itself.  Halo samples are
// recomputed by neighbouring tiles from the same source, so tiles are
// independent and no second launch or grid barrier is needed.
//
// PlaneBuffer semantics are kept: reads clamp to the plane, writes outside
// the plane are dropped (loop_filter.rs:128-158).
#pragma once

#include "av1_tables.h"
#include "common.hpp"

namespace oxav1 {

constexpr int kLfTile = 64;
constexpr int kLfHalo = 12;
constexpr int kLfRegion = kLfTile + 2 * kLfHalo;   // 88
constexpr int kLfStride = kLfRegion + 2;           // LDS row pitch: 45 dwords of u16, odd, so
                                                   // lanes walking down a column hit distinct banks
constexpr int kLfEdges = kLfTile / 4 + 3;          // edges at X0-4 .. X0+TW+4
constexpr int kLfCellRows = kLfRegion / 4;         // 22

__constant__ uint8_t c_tx_w[19] = {4, 8, 16, 32, 64, 4, 8, 8, 16, 16, 32, 32, 64, 4, 16, 8, 32, 16, 64};
__constant__ uint8_t c_tx_h[19] = {4, 8, 16, 32, 64, 8, 4, 16, 8, 32, 16, 64, 32, 16, 4, 32, 8, 64, 16};
__constant__ uint8_t c_num4x4_w[22] = {1, 1, 2, 2, 2, 4, 4, 4, 8, 8, 8, 16, 16, 16, 32, 32, 1, 4, 2, 8, 4, 16};
__constant__ uint8_t c_num4x4_h[22] = {1, 2, 1, 2, 4, 2, 4, 8, 4, 8, 16, 8, 16, 32, 16, 32, 4, 1, 8, 2, 16, 4};
__constant__ uint8_t c_subsampled[88];  // filled from kSubsampledSize at init

// adaptive_filter_strength_selection (loop_filter.rs:527-572) for the cell's
// packed word; returns lvl.
__device__ __forceinline__ uint32_t lf_level_at(const FrameJob &J, uint32_t row, uint32_t col,
                                                uint32_t plane, uint32_t pass) {
    const size_t k = (size_t)row * J.mi_cols + col;
    const uint32_t w = gread(J.lf_mi, k);
    const uint32_t seg = w >> OXAV1_MI_SEG_SHIFT;
    const uint32_t ref = (w >> OXAV1_MI_REF_SHIFT) & 15u;
    const uint32_t mode = (w >> OXAV1_MI_MODE_SHIFT) & 31u;
    const uint32_t mode_type = (mode >= 14 && mode != 16 && mode != 24) ? 1u : 0u;
    const uint32_t i = plane == 0 ? pass : plane + 1;
    const uint32_t dli = J.lf_delta_multi ? i : 0u;
    const int32_t dlf = J.delta_lf ? (int32_t)gread(J.delta_lf, k * 4 + dli) : 0;
    int32_t lvl = clip3(0, 63, dlf + (int32_t)J.lf_level[i]);
    const uint32_t feature = 1 + i;
    if (seg < 8 && J.seg_active[seg][feature]) lvl = clip3(0, 63, (int32_t)J.seg_data[seg][feature] + lvl);
    if (J.lf_delta_enabled) {
        const int32_t n_shift = lvl >> 5;
        if (ref == 0) {
            lvl += (int32_t)J.lf_ref_deltas[0] * (1 << n_shift);
        } else {
            const uint32_t ri = ref < 7 ? ref : 7;
            lvl += (int32_t)J.lf_ref_deltas[ri] * (1 << n_shift);
            lvl += (int32_t)J.lf_mode_deltas[mode_type] * (1 << n_shift);
        }
        lvl = clip3(0, 63, lvl);
    }
    return (uint32_t)lvl;
}

// loop_filter_edge (loop_filter.rs:302-397) up to the strength: returns 0 for
// "no filtering", else (filter_size | lvl << 8).  (by, bx) index the plane's
// 4x4 blocks, i.e. the loop cursor row = by << sy, col = bx << sx
// (loop_filter.rs:282-292).
__device__ __forceinline__ uint32_t lf_decide(const FrameJob &J, uint32_t plane, uint32_t pass, int32_t by,
                              int32_t bx) {
    if (by < 0 || bx < 0) return 0;
    const uint32_t sx = plane ? J.sx : 0, sy = plane ? J.sy : 0;
    uint32_t row = (uint32_t)by << sy, col = (uint32_t)bx << sx;
    if (row >= J.mi_rows || col >= J.mi_cols) return 0;  // not visited by the walk
    const uint32_t x = col * 4, y = row * 4;
    row |= sy;
    col |= sx;
    if (x >= J.fw || y >= J.fh) return 0;
    if (pass == 0 && x == 0) return 0;
    if (pass == 1 && y == 0) return 0;
    const uint32_t xp = x >> sx, yp = y >> sy;
    const int32_t prev_row = (int32_t)row - ((pass == 1 ? 1 : 0) << sy);
    const int32_t prev_col = (int32_t)col - ((pass == 0 ? 1 : 0) << sx);
    if (prev_row < 0 || prev_col < 0) return 0;
    if (row >= J.mi_rows || col >= J.mi_cols) return 0;
    const uint32_t w = gread(J.lf_mi, (size_t)row * J.mi_cols + col);
    const uint32_t mi_size = (w >> OXAV1_MI_SIZE_SHIFT) & 31u;
    const uint32_t tcols = J.tx_cols[plane];
    const uint32_t tx_sz = gread(J.tx[plane], (size_t)(row >> sy) * tcols + (col >> sx));
    const uint32_t ptx = gread(J.tx[plane], (size_t)((uint32_t)prev_row >> sy) * tcols + ((uint32_t)prev_col >> sx));
    // planeSize = get_plane_residual_size(MiSize, plane).unwrap_or(MiSize)
    uint32_t plane_size = mi_size;
    if (mi_size < 22) {
        const uint32_t subx = plane ? (J.sx != 0) : 0, suby = plane ? (J.sy != 0) : 0;
        const uint32_t raw = c_subsampled[mi_size * 4 + subx * 2 + suby];
        if (raw != OXAV1_BLOCK_INVALID) plane_size = raw;
    }
    const bool skip = (w >> OXAV1_MI_SKIP_BIT) & 1u;
    const bool is_intra = (w >> OXAV1_MI_INTRA_BIT) & 1u;
    bool block_edge, tx_edge;
    uint32_t fs;
    if (pass == 0) {
        const uint32_t bw = plane_size < 22 ? (uint32_t)c_num4x4_w[plane_size] * 4 : 4u;
        block_edge = (xp % bw) == 0;
        tx_edge = (xp % c_tx_w[tx_sz]) == 0;
        fs = min((uint32_t)c_tx_w[tx_sz], (uint32_t)c_tx_w[ptx]);
    } else {
        const uint32_t bh = plane_size < 22 ? (uint32_t)c_num4x4_h[plane_size] * 4 : 4u;
        block_edge = (yp % bh) == 0;
        tx_edge = (yp % c_tx_h[tx_sz]) == 0;
        fs = min((uint32_t)c_tx_h[tx_sz], (uint32_t)c_tx_h[ptx]);
    }
    fs = min(fs, plane == 0 ? 16u : 8u);
    const bool apply = tx_edge && (block_edge || !skip || is_intra);
    if (!apply) return 0;
    uint32_t lvl = lf_level_at(J, row, col, plane, pass);
    if (lvl == 0) lvl = lf_level_at(J, (uint32_t)prev_row, (uint32_t)prev_col, plane, pass);
    if (lvl == 0) return 0;
    return fs | (lvl << 8);
}

// LDS view of the tile region with the reference's clamped reads / dropped
// writes.  Region origin (ry0, rx0) may be negative; only in-plane entries are
// loaded and every clamped coordinate lands inside the region.  EDGE = false
// when the whole region lies inside the plane: every tap of every edge the
// tile filters then stays inside the plane and the region, so the clamps and
// the write test are no-ops and are skipped.
template <typename L, bool EDGE>
struct LfRegion {
    L *s;
    int32_t ry0, rx0, rows, cols;
    __device__ __forceinline__ int idx(int32_t y, int32_t x) const {
        if (EDGE) {
            y = clip3(0, rows - 1, y);
            x = clip3(0, cols - 1, x);
        }
        return (y - ry0) * kLfStride + (x - rx0);
    }
    __device__ __forceinline__ int32_t get(int32_t y, int32_t x) const { return (int32_t)s[idx(y, x)]; }
    __device__ __forceinline__ void set(int32_t y, int32_t x, int32_t v) const {
        if (EDGE && (y < 0 || x < 0 || y >= rows || x >= cols)) return;
        s[(y - ry0) * kLfStride + (x - rx0)] = (L)v;
    }
};

// lo = -(1 << (bd - 1)) < hi = (1 << (bd - 1)) - 1: the clamp is the median of
// three, one v_med3_i32
__device__ __forceinline__ int32_t f4clamp(int32_t v, int32_t lo, int32_t hi) {
    int32_t r;
    asm("v_med3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(v), "v"(lo), "v"(hi));
    return r;
}

// wide_filter (loop_filter.rs:789-822) with N = n, N2 = n2, LOG2 = log2Size
// fixed at compile time: output i in [-N, N) is
// Round2(sum_{j=-N..N} v[Clip3(-(N+1), N, i+j)] * (|j| <= N2 ? 2 : 1), LOG2).
template <int N, int N2, int LOG2, typename Reg>
__device__ __forceinline__ void lf_wide(const Reg &R, const int32_t *v, int32_t x, int32_t y, int32_t dx,
                                        int32_t dy) {
    int32_t out[2 * N];
#pragma unroll
    for (int i = -N; i < N; i++) {
        int32_t t = 0;
#pragma unroll
        for (int j = -N; j <= N; j++) {
            const int pidx = i + j < -(N + 1) ? -(N + 1) : (i + j > N ? N : i + j);
            t += v[pidx + 7] * ((j >= -N2 && j <= N2) ? 2 : 1);
        }
        out[i + N] = (t + (1 << (LOG2 - 1))) >> LOG2;
    }
#pragma unroll
    for (int i = -N; i < N; i++) R.set(y + i * dy, x + i * dx, out[i + N]);
}

// sample_filtering + filter_mask + narrow/wide filter (loop_filter.rs:582-822)
// for one sample position of one edge.
template <typename L, bool EDGE>
__device__ __forceinline__ void lf_filter_sample(const LfRegion<L, EDGE> &R, int32_t x, int32_t y, int32_t dx, int32_t dy,
                                 uint32_t plane, uint32_t fs, uint32_t lvl, uint32_t sharp, uint32_t bd) {
    // strength (loop_filter.rs:502-523)
    const int32_t shift = sharp > 4 ? 2 : (sharp > 0 ? 1 : 0);
    const int32_t limit = sharp > 0 ? clip3(1, 9 - (int32_t)sharp, (int32_t)lvl >> shift)
                                    : max((int32_t)lvl >> shift, 1);
    const int32_t blimit = 2 * ((int32_t)lvl + 2) + limit;
    const int32_t thresh = (int32_t)lvl >> 4;
    int32_t q[7], p[7];
#pragma unroll
    for (int k = 0; k < 7; k++) {
        q[k] = R.get(y + k * dy, x + k * dx);
        p[k] = R.get(y - (k + 1) * dy, x - (k + 1) * dx);
    }
    const uint32_t bds = bd - 8;
    const int32_t thresh_bd = thresh << bds, limit_bd = limit << bds, blimit_bd = blimit << bds;
    const bool hev = abs(p[1] - p[0]) > thresh_bd || abs(q[1] - q[0]) > thresh_bd;
    const uint32_t flen = fs == 4 ? 4 : (plane != 0 ? 6 : (fs == 8 ? 8 : 16));
    bool m = abs(p[1] - p[0]) > limit_bd;
    m |= abs(q[1] - q[0]) > limit_bd;
    m |= abs(p[0] - q[0]) * 2 + abs(p[1] - q[1]) / 2 > blimit_bd;
    if (flen >= 6) {
        m |= abs(p[2] - p[1]) > limit_bd;
        m |= abs(q[2] - q[1]) > limit_bd;
    }
    if (flen >= 8) {
        m |= abs(p[3] - p[2]) > limit_bd;
        m |= abs(q[3] - q[2]) > limit_bd;
    }
    if (m) return;  // filterMask == 0
    const int32_t one = 1 << bds;
    bool flat = false, flat2 = false;
    if (fs >= 8) {
        bool f = abs(p[1] - p[0]) > one || abs(q[1] - q[0]) > one || abs(p[2] - p[0]) > one ||
                 abs(q[2] - q[0]) > one;
        if (flen >= 8) f = f || abs(p[3] - p[0]) > one || abs(q[3] - q[0]) > one;
        flat = !f;
    }
    if (fs >= 16) {
        bool f = abs(p[6] - p[0]) > one || abs(q[6] - q[0]) > one || abs(p[5] - p[0]) > one ||
                 abs(q[5] - q[0]) > one || abs(p[4] - p[0]) > one || abs(q[4] - q[0]) > one;
        flat2 = !f;
    }
    if (fs == 4 || !flat) {
        // narrow_filter (loop_filter.rs:743-782)
        const int32_t lo = -(1 << (bd - 1)), hi = (1 << (bd - 1)) - 1;
        const int32_t half = 0x80 << bds;
        const int32_t ps1 = p[1] - half, ps0 = p[0] - half, qs0 = q[0] - half, qs1 = q[1] - half;
        int32_t filt = hev ? f4clamp(ps1 - qs1, lo, hi) : 0;
        filt = f4clamp(filt + 3 * (qs0 - ps0), lo, hi);
        const int32_t f1 = f4clamp(filt + 4, lo, hi) >> 3;
        const int32_t f2 = f4clamp(filt + 3, lo, hi) >> 3;
        R.set(y, x, f4clamp(qs0 - f1, lo, hi) + half);
        R.set(y - dy, x - dx, f4clamp(ps0 + f2, lo, hi) + half);
        if (!hev) {
            const int32_t f = (f1 + 1) >> 1;
            R.set(y + dy, x + dx, f4clamp(qs1 - f, lo, hi) + half);
            R.set(y - 2 * dy, x - 2 * dx, f4clamp(ps1 + f, lo, hi) + half);
        }
        return;
    }
    // wide_filter (loop_filter.rs:789-822): 13-tap (flat2, filter size 16),
    // 7-tap luma or 5-tap chroma (filter size >= 8)
    int32_t v[14];  // v[t + 7] for t = -7..6: p6..p0 then q0..q6
#pragma unroll
    for (int k = 0; k < 7; k++) {
        v[6 - k] = p[k];
        v[7 + k] = q[k];
    }
    if (fs >= 16 && flat2)
        lf_wide<6, 1, 4>(R, v, x, y, dx, dy);
    else if (plane == 0)
        lf_wide<3, 0, 3>(R, v, x, y, dx, dy);
    else
        lf_wide<2, 1, 3>(R, v, x, y, dx, dy);
}

// Passes 0 and 1 over the loaded region (see lf_tile).
template <typename L, bool EDGE>
__device__ __forceinline__ void lf_passes(const LfRegion<L, EDGE> &R, int32_t X0, int32_t Y0, uint32_t plane,
                          const uint16_t *dec0, const uint16_t *dec1, uint32_t sharp, uint32_t bd) {
    const int tid = threadIdx.x, nt = blockDim.x;
    // pass 0: vertical edges x every loaded row; consecutive lanes take
    // consecutive rows of one edge (one decision per 4 lanes)
    // (loops run a workgroup-uniform trip count; the work is predicated)
    for (int base = 0; base < kLfRegion * kLfEdges; base += nt) {
        const int i = base + tid;
        const int e = i / kLfRegion, ry = i - e * kLfRegion;
        const int32_t y = R.ry0 + ry;
        const bool rowok = !EDGE || (y >= 0 && y < R.rows);
        const uint32_t d = (i < kLfRegion * kLfEdges && rowok) ? dec0[(ry >> 2) * kLfEdges + e] : 0u;
        if (d) lf_filter_sample(R, X0 - 4 + 4 * e, y, 1, 0, plane, d & 0xffu, d >> 8, sharp, bd);
    }
    __syncthreads();
    // pass 1: horizontal edges x tile columns (consecutive lanes, consecutive columns)
    for (int base = 0; base < kLfEdges * kLfTile; base += nt) {
        const int i = base + tid;
        const int er = i / kLfTile, c = i - er * kLfTile;
        const int32_t x = X0 + c;
        const bool colok = !EDGE || x < R.cols;
        const uint32_t d = (i < kLfEdges * kLfTile && colok) ? dec1[er * (kLfTile / 4) + (c >> 2)] : 0u;
        if (d) lf_filter_sample(R, x, Y0 - 4 + 4 * er, 0, 1, plane, d & 0xffu, d >> 8, sharp, bd);
    }
}

// One 64x64 tile of one plane: src (T) -> dst (T).
template <typename T>
__device__ __forceinline__ void lf_tile(const FrameJob &J, uint32_t plane, uint32_t ty, uint32_t tx,
                        const G<T> *__restrict__ src, uint32_t spitch, G<T> *__restrict__ dst,
                        uint32_t dpitch, typename LdsOf<T>::type *sm, uint16_t *dec0, uint16_t *dec1) {
    using L = typename LdsOf<T>::type;
    const int32_t rows = (int32_t)J.pad_rows[plane], cols = (int32_t)J.pad_cols[plane];
    const int32_t Y0 = (int32_t)ty * kLfTile, X0 = (int32_t)tx * kLfTile;
    const int32_t ry0 = Y0 - kLfHalo, rx0 = X0 - kLfHalo;
    const bool interior = ry0 >= 0 && rx0 >= 0 && ry0 + kLfRegion <= rows && rx0 + kLfRegion <= cols;
    const int tid = threadIdx.x, nt = blockDim.x;

    // 1. load the in-plane part of the region
    // (every loop below runs a workgroup-uniform trip count; work is predicated)
    // In rounds of kRound loads per thread, every load of a round issued (at
    // a clamped, in-plane index) before its LDS stores, so latencies overlap.
    {
        constexpr int kN = kLfRegion * kLfRegion, kRound = 8;
        static_assert(4 * kRound * 256 >= kN, "region fits four rounds");
#pragma unroll 1
        for (int r0 = 0; r0 < 4 * kRound; r0 += kRound) {
            L v[kRound];
#pragma unroll
            for (int k = 0; k < kRound; k++) {
                const int i = min((r0 + k) * nt + tid, kN - 1);
                const int ry = i / kLfRegion, rx = i - ry * kLfRegion;
                const int32_t y = clip3(0, rows - 1, ry0 + ry), x = clip3(0, cols - 1, rx0 + rx);
                v[k] = (L)src[(size_t)y * spitch + x];
            }
#pragma unroll
            for (int k = 0; k < kRound; k++) {
                const int i = (r0 + k) * nt + tid;
                const int ry = i / kLfRegion, rx = i - ry * kLfRegion;
                const int32_t y = ry0 + ry, x = rx0 + rx;
                if (i < kN && (interior || (y >= 0 && y < rows && x >= 0 && x < cols))) sm[ry * kLfStride + rx] = v[k];
            }
        }
    }
    // 2. edge decisions: pass 0 for region cell rows x edges X0-4..X0+TW+4,
    //    pass 1 for edge rows Y0-4..Y0+TH+4 x tile cell columns.
    const bool lf_on = J.apply_lf && J.lf_plane_on[plane];
    for (int base = 0; base < kLfCellRows * kLfEdges; base += nt) {
        const int i = base + tid;
        const int cr = i / kLfEdges, e = i - cr * kLfEdges;
        const int32_t by = (ry0 >> 2) + cr, bx = ((X0 - 4) >> 2) + e;
        if (i < kLfCellRows * kLfEdges) dec0[i] = lf_on ? (uint16_t)lf_decide(J, plane, 0, by, bx) : 0;
    }
    for (int base = 0; base < kLfEdges * (kLfTile / 4); base += nt) {
        const int i = base + tid;
        const int er = i / (kLfTile / 4), c = i - er * (kLfTile / 4);
        const int32_t by = ((Y0 - 4) >> 2) + er, bx = (X0 >> 2) + c;
        if (i < kLfEdges * (kLfTile / 4)) dec1[i] = lf_on ? (uint16_t)lf_decide(J, plane, 1, by, bx) : 0;
    }
    __syncthreads();
    const uint32_t sharp = J.lf_sharp, bd = J.bd;
    // 3./4. pass 0 and pass 1
    if (interior)
        lf_passes(LfRegion<L, false>{sm, ry0, rx0, rows, cols}, X0, Y0, plane, dec0, dec1, sharp, bd);
    else
        lf_passes(LfRegion<L, true>{sm, ry0, rx0, rows, cols}, X0, Y0, plane, dec0, dec1, sharp, bd);
    __syncthreads();
    // 5. store the tile
    for (int base = 0; base < kLfTile * kLfTile; base += nt) {
        const int i = base + tid;
        const int yy = i >> 6, xx = i & 63;
        const int32_t y = Y0 + yy, x = X0 + xx;
        if (i < kLfTile * kLfTile && y < rows && x < cols)
            dst[(size_t)y * dpitch + x] = (T)sm[(yy + kLfHalo) * kLfStride + (xx + kLfHalo)];
    }
}

}  // namespace oxav1
