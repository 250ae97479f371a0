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
// Work inside a tile:
// 1. the region is loaded four samples per lane and row (one vector load),
//    rows clamped to the plane;
// 2. one thread per 4x4 cell of the region decides both passes' edges of the
//    cell (loop_filter_edge up to the strength); edges that filter are
//    pushed, as (cell, level) entries, onto per-pass lists bucketed by filter
//    size (4, 8, 16), with a wave ballot and one LDS atomic per wave;
// 3. each pass walks its three lists, four lanes per entry (the four sample
//    rows or columns of the cell's edge), each bucket with its filter size
//    fixed at compile time: a wave runs one mask and filter path and reads
//    only the taps that filter size uses;
// 4. the tile is stored four samples per lane.
// On typical content most edge samples fail the filter mask, so the cost is
// in the loads, the decisions and the masks; edges with nothing to filter
// take no lanes at all.
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
constexpr int kLfEdges = kLfTile / 4 + 3;          // edge columns X0-4 .. X0+TW+4
constexpr int kLfCellRows = kLfRegion / 4;         // 22
constexpr int kLfCells = kLfCellRows * kLfEdges;   // 418 cells decided per tile
constexpr int kLfCells1 = kLfEdges * (kLfTile / 4);  // 304 of them carry a pass-1 edge
constexpr int kLfGroups = kLfRegion / 4;           // 22 four-sample groups per region row
constexpr int kLfLoadRows = 11;                    // region rows per load round (22 x 11 = 242 lanes)

__constant__ uint8_t c_tx_w[19] = {4, 8, 16, 32, 64, 4, 8, 8, 16, 16, 32, 32, 64, 4, 16, 8, 32, 16, 64};
__constant__ uint8_t c_tx_h[19] = {4, 8, 16, 32, 64, 8, 4, 16, 8, 32, 16, 64, 32, 16, 4, 32, 8, 64, 16};
__constant__ uint8_t c_num4x4_w[22] = {1, 1, 2, 2, 2, 4, 4, 4, 8, 8, 8, 16, 16, 16, 32, 32, 1, 4, 2, 8, 4, 16};
__constant__ uint8_t c_num4x4_h[22] = {1, 2, 1, 2, 4, 2, 4, 8, 4, 8, 16, 8, 16, 32, 16, 32, 4, 1, 8, 2, 16, 4};
__constant__ uint8_t c_subsampled[88];  // filled from kSubsampledSize at init

// log2 of the 19 transform sizes' width / height in samples, three bits per
// size (TX_4X4 = 0 .. TX_64X16 = 18), held in registers instead of a table
template <size_t N>
constexpr uint64_t lf_pack3(const uint8_t (&v)[N]) {
    uint64_t r = 0;
    for (size_t i = 0; i < N; i++) r |= (uint64_t)v[i] << (3 * i);
    return r;
}
constexpr uint8_t kTxWLog2[19] = {2, 3, 4, 5, 6, 2, 3, 3, 4, 4, 5, 5, 6, 2, 4, 3, 5, 4, 6};
constexpr uint8_t kTxHLog2[19] = {2, 3, 4, 5, 6, 3, 2, 4, 3, 5, 4, 6, 5, 4, 2, 5, 3, 6, 4};
constexpr uint64_t kTxWLog2P = lf_pack3(kTxWLog2);
constexpr uint64_t kTxHLog2P = lf_pack3(kTxHLog2);
__device__ __forceinline__ uint32_t lf_txl(uint64_t packed, uint32_t tx) { return (uint32_t)(packed >> (3 * tx)) & 7u; }

// LDS of one deblocking work item besides the sample region.
struct LfShared {
    uint16_t list0[3][kLfCells];   // pass-0 entries per filter-size bucket (4, 8, 16)
    uint16_t list1[3][kLfCells1];  // pass-1 entries
    uint32_t cnt[2][3];            // entries per (pass, bucket)
    uint32_t str[64];              // per level: limit << bds | blimit << bds << 12 | thresh << bds << 24
    int16_t segd[2][8];            // seg_data[seg][1 + i] of each pass's level index i
    uint8_t dims[32];              // per mi_size: log2 plane block width | log2 height << 4
};

// A list entry: edge column e (5 bits) | cell row cr (5 bits) | level (6 bits).
__device__ __forceinline__ uint16_t lf_entry(uint32_t e, uint32_t cr, uint32_t lvl) {
    return (uint16_t)(e | cr << 5 | lvl << 10);
}

// Values every lane of the tile shares, in scalar registers.
struct LfUni {
    const G<uint32_t> *mi;
    const G<uint8_t> *tx;
    const G<int8_t> *dlf;   // null without per-block deltas
    uint32_t sx, sy, mi_rows, mi_cols, fw, fh, tcols, cap;  // cap: log2 of the largest filter size
    uint32_t base[2];       // lf_level[i] of pass 0 / 1
    uint32_t dli[2];        // delta_lf index of pass 0 / 1
    uint32_t segm[2];       // bit s: segment s has the pass's level feature active
    uint64_t refd;          // lf_ref_deltas, one byte each
    int32_t md0, md1;       // lf_mode_deltas
    uint32_t delta_en;
};

__device__ __forceinline__ int32_t lf_sbyte(uint64_t packed, uint32_t i) {
    return (int32_t)(int8_t)(uint8_t)(packed >> (8 * i));
}

// adaptive_filter_strength_selection (loop_filter.rs:527-572) of pass P for
// the mi cell at packed word w, mi index k.
template <int P>
__device__ __forceinline__ uint32_t lf_level(const LfUni &U, const int16_t *segd, uint32_t w, uint32_t k) {
    const uint32_t seg = w >> OXAV1_MI_SEG_SHIFT;
    const uint32_t ref = (w >> OXAV1_MI_REF_SHIFT) & 15u;
    const uint32_t mode = (w >> OXAV1_MI_MODE_SHIFT) & 31u;
    const int32_t dlf = U.dlf ? (int32_t)U.dlf[k * 4 + U.dli[P]] : 0;
    int32_t lvl = clip3(0, 63, dlf + (int32_t)U.base[P]);
    if (seg < 8 && ((U.segm[P] >> seg) & 1u)) lvl = clip3(0, 63, (int32_t)segd[seg] + lvl);
    if (U.delta_en) {
        const int32_t n_shift = lvl >> 5;
        int32_t d = lf_sbyte(U.refd, ref < 7 ? ref : 7u);  // ref 0: lf_ref_deltas[0], no mode delta
        if (ref != 0) d += (mode >= 14 && mode != 16 && mode != 24) ? U.md1 : U.md0;
        lvl = clip3(0, 63, lvl + d * (1 << n_shift));
    }
    return (uint32_t)lvl;
}

// loop_filter_edge (loop_filter.rs:302-397) up to the strength, for both
// passes of the plane's 4x4 cell (by, bx): d0 / d1 = 0 for "no filtering",
// else lvl | (log2 filter size - 2) << 6.  The loop cursor of the reference
// is row = by << sy, col = bx << sx (loop_filter.rs:282-292).
template <bool W0, bool W1>
__device__ __forceinline__ void lf_decide(const LfUni &U, const LfShared &S, int32_t by, int32_t bx, bool want0,
                                          bool want1, uint32_t &d0, uint32_t &d1) {
    d0 = d1 = 0;
    if (by < 0 || bx < 0) return;
    uint32_t row = (uint32_t)by << U.sy, col = (uint32_t)bx << U.sx;
    if (row >= U.mi_rows || col >= U.mi_cols) return;  // not visited by the walk
    const uint32_t x = col * 4, y = row * 4;
    row |= U.sy;
    col |= U.sx;
    if (x >= U.fw || y >= U.fh) return;
    if (row >= U.mi_rows || col >= U.mi_cols) return;
    const uint32_t k = row * U.mi_cols + col;
    const uint32_t w = U.mi[k];
    const uint32_t t = (uint32_t)by * U.tcols + (uint32_t)bx;
    const uint32_t tx_sz = U.tx[t];
    // planeSize = get_plane_residual_size(MiSize, plane).unwrap_or(MiSize)
    const uint32_t dm = S.dims[w & 31u];
    const bool skip = (w >> OXAV1_MI_SKIP_BIT) & 1u;
    const bool is_intra = (w >> OXAV1_MI_INTRA_BIT) & 1u;
    const bool open = !skip || is_intra;  // block edge not required
    // x == 0 / y == 0 edges are not filtered; then prev_col / prev_row >= 0
    if (W0 && want0 && x != 0) {
        const uint32_t xp = (uint32_t)bx * 4;
        const uint32_t tw = lf_txl(kTxWLog2P, tx_sz);
        const bool tx_edge = (xp & ((1u << tw) - 1)) == 0;
        const bool block_edge = (xp & ((1u << (dm & 15u)) - 1)) == 0;
        if (tx_edge && (block_edge || open)) {
            const uint32_t ptx = U.tx[t - 1];
            const uint32_t fl = min(min(tw, lf_txl(kTxWLog2P, ptx)), U.cap);
            uint32_t lvl = lf_level<0>(U, S.segd[0], w, k);
            if (lvl == 0) {
                const uint32_t kp = k - (1u << U.sx);
                lvl = lf_level<0>(U, S.segd[0], U.mi[kp], kp);
            }
            if (lvl) d0 = lvl | (fl - 2) << 6;
        }
    }
    if (W1 && want1 && y != 0) {
        const uint32_t yp = (uint32_t)by * 4;
        const uint32_t th = lf_txl(kTxHLog2P, tx_sz);
        const bool tx_edge = (yp & ((1u << th) - 1)) == 0;
        const bool block_edge = (yp & ((1u << (dm >> 4)) - 1)) == 0;
        if (tx_edge && (block_edge || open)) {
            const uint32_t ptx = U.tx[t - U.tcols];
            const uint32_t fl = min(min(th, lf_txl(kTxHLog2P, ptx)), U.cap);
            uint32_t lvl = lf_level<1>(U, S.segd[1], w, k);
            if (lvl == 0) {
                const uint32_t kp = k - (U.mi_cols << U.sy);
                lvl = lf_level<1>(U, S.segd[1], U.mi[kp], kp);
            }
            if (lvl) d1 = lvl | (fl - 2) << 6;
        }
    }
}

// Appends ent to list (at the slot counter *cnt hands out) on every lane
// where pred holds: one ballot, one LDS atomic per wave.  Called with the
// whole wave active.
__device__ __forceinline__ void lf_push(uint32_t *cnt, uint16_t *list, bool pred, uint16_t ent) {
    const uint64_t m = __ballot(pred);
    if (m == 0) return;
    uint32_t base = 0;
    if (__lane_id() == 0) base = atomicAdd(cnt, (uint32_t)__popcll(m));
    base = __builtin_amdgcn_readfirstlane(base);
    const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
    if (pred) list[base + rank] = ent;
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
    // (y, x) inside the region: 0 <= y - ry0 < kLfRegion (said to the
    // compiler, which then forms the row offset on the 24-bit multiplier)
    __device__ __forceinline__ int at(int32_t y, int32_t x) const {
        const int32_t dy = y - ry0;
        __builtin_assume((uint32_t)dy < (uint32_t)kLfRegion);
        return dy * kLfStride + (x - rx0);
    }
    __device__ __forceinline__ int idx(int32_t y, int32_t x) const {
        if (EDGE) {
            y = clip3(0, rows - 1, y);
            x = clip3(0, cols - 1, x);
        }
        return at(y, x);
    }
    __device__ __forceinline__ int32_t get(int32_t y, int32_t x) const { return (int32_t)s[idx(y, x)]; }
    __device__ __forceinline__ void set(int32_t y, int32_t x, int32_t v) const {
        if (EDGE && (y < 0 || x < 0 || y >= rows || x >= cols)) return;
        s[at(y, x)] = (L)v;
    }
};

// lo = -(1 << (bd - 1)) < hi = (1 << (bd - 1)) - 1: the clamp is the median of
// three, one v_med3_i32
__device__ __forceinline__ int32_t f4clamp(int32_t v, int32_t lo, int32_t hi) {
    int32_t r;
    asm("v_med3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(v), "v"(lo), "v"(hi));
    return r;
}

// |a - b| of two samples: one v_sad_u16 for u8/u16 frames (samples < 2^16);
// the int32 path keeps the reference's i32 arithmetic.
template <typename L>
__device__ __forceinline__ int32_t lf_ad(int32_t a, int32_t b) {
    if constexpr (std::is_same<L, int32_t>::value)
        return abs(a - b);
    else
        return (int32_t)__builtin_amdgcn_sad_u16((uint32_t)a, (uint32_t)b, 0u);
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
// for one sample position of one edge whose filter size is 1 << FL (filter
// length 4, 6 (chroma 8), 8 or 16), (DX, DY) the step across the edge.
// str: limit_bd | blimit_bd << 12 | thresh_bd << 24.
template <int FL, bool LUMA, int DX, int DY, typename L, bool EDGE>
__device__ __forceinline__ void lf_filter_sample(const LfRegion<L, EDGE> &R, int32_t x, int32_t y, uint32_t str,
                                                 uint32_t bd) {
    constexpr int FS = 1 << FL;
    constexpr int FLEN = FS == 4 ? 4 : (!LUMA ? 6 : (FS == 8 ? 8 : 16));
    constexpr int NT = FS == 16 ? 7 : FLEN / 2;  // taps read on each side
    const int32_t limit_bd = (int32_t)(str & 0xfffu), blimit_bd = (int32_t)((str >> 12) & 0xfffu);
    const int32_t thresh_bd = (int32_t)(str >> 24);
    const uint32_t bds = bd - 8;
    int32_t q[7], p[7];
#pragma unroll
    for (int k = 0; k < NT; k++) {
        q[k] = R.get(y + k * DY, x + k * DX);
        p[k] = R.get(y - (k + 1) * DY, x - (k + 1) * DX);
    }
    const int32_t a10 = max(lf_ad<L>(p[1], p[0]), lf_ad<L>(q[1], q[0]));
    const bool hev = a10 > thresh_bd;
    int32_t am = a10;
    if (FLEN >= 6) am = max(am, max(lf_ad<L>(p[2], p[1]), lf_ad<L>(q[2], q[1])));
    if (FLEN >= 8) am = max(am, max(lf_ad<L>(p[3], p[2]), lf_ad<L>(q[3], q[2])));
    bool m = am > limit_bd;
    if constexpr (std::is_same<L, int32_t>::value)
        m |= abs(p[0] - q[0]) * 2 + abs(p[1] - q[1]) / 2 > blimit_bd;
    else
        m |= (uint32_t)lf_ad<L>(p[0], q[0]) * 2 + ((uint32_t)lf_ad<L>(p[1], q[1]) >> 1) > (uint32_t)blimit_bd;
    if (m) return;  // filterMask == 0
    const int32_t one = 1 << bds;
    bool flat = false, flat2 = false;
    if (FS >= 8) {
        int32_t f = max(a10, max(lf_ad<L>(p[2], p[0]), lf_ad<L>(q[2], q[0])));
        if (FLEN >= 8) f = max(f, max(lf_ad<L>(p[3], p[0]), lf_ad<L>(q[3], q[0])));
        flat = !(f > one);
    }
    if (FS >= 16) {
        const int32_t f = max(max(max(lf_ad<L>(p[6], p[0]), lf_ad<L>(q[6], q[0])),
                                  max(lf_ad<L>(p[5], p[0]), lf_ad<L>(q[5], q[0]))),
                              max(lf_ad<L>(p[4], p[0]), lf_ad<L>(q[4], q[0])));
        flat2 = !(f > one);
    }
    if (FS == 4 || !flat) {
        // narrow_filter (loop_filter.rs:743-782)
        const int32_t lo = -(1 << (bd - 1)), hi = (1 << (bd - 1)) - 1;
        const int32_t half = 0x80 << bds;
        const int32_t ps1 = p[1] - half, ps0 = p[0] - half, qs0 = q[0] - half, qs1 = q[1] - half;
        int32_t filt = hev ? f4clamp(ps1 - qs1, lo, hi) : 0;
        // 16-bit samples: |qs0 - ps0| < 2^16, a 24-bit product
        const int32_t d3 = sizeof(L) == 2 ? mul_i24k<3>(qs0 - ps0) : 3 * (qs0 - ps0);
        filt = f4clamp(filt + d3, lo, hi);
        const int32_t f1 = f4clamp(filt + 4, lo, hi) >> 3;
        const int32_t f2 = f4clamp(filt + 3, lo, hi) >> 3;
        R.set(y, x, f4clamp(qs0 - f1, lo, hi) + half);
        R.set(y - DY, x - DX, f4clamp(ps0 + f2, lo, hi) + half);
        if (!hev) {
            const int32_t f = (f1 + 1) >> 1;
            R.set(y + DY, x + DX, f4clamp(qs1 - f, lo, hi) + half);
            R.set(y - 2 * DY, x - 2 * DX, f4clamp(ps1 + f, lo, hi) + half);
        }
        return;
    }
    // wide_filter (loop_filter.rs:789-822): 13-tap (flat2, filter size 16),
    // 7-tap luma or 5-tap chroma (filter size >= 8)
    int32_t v[14];  // v[t + 7] for t = -7..6: p6..p0 then q0..q6
#pragma unroll
    for (int k = 0; k < 7; k++) {
        v[6 - k] = k < NT ? p[k] : 0;
        v[7 + k] = k < NT ? q[k] : 0;
    }
    if (FS >= 16 && flat2)
        lf_wide<6, 1, 4>(R, v, x, y, DX, DY);
    else if (LUMA)
        lf_wide<3, 0, 3>(R, v, x, y, DX, DY);
    else
        lf_wide<2, 1, 3>(R, v, x, y, DX, DY);
}

// One bucket of one pass: four lanes per entry, one per sample row (pass 0)
// or column (pass 1) of the entry's 4-sample edge.  n is workgroup-uniform.
template <int PASS, int FL, bool LUMA, typename L, bool EDGE>
__device__ __forceinline__ void lf_bucket(const LfRegion<L, EDGE> &R, const uint16_t *list, uint32_t n, int32_t X0,
                                          const LfShared &S, uint32_t bd) {
    const uint32_t lanes = n * 4;
    for (uint32_t base = 0; base < lanes; base += blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        if (i < lanes) {
            const uint32_t ent = list[i >> 2];
            const uint32_t e = ent & 31u, cr = (ent >> 5) & 31u, lvl = ent >> 10;
            int32_t x = X0 - 4 + 4 * (int32_t)e, y = R.ry0 + 4 * (int32_t)cr;
            bool ok;
            if (PASS == 0) {
                y += (int32_t)(i & 3u);
                ok = !EDGE || (y >= 0 && y < R.rows);
            } else {
                x += (int32_t)(i & 3u);
                ok = !EDGE || x < R.cols;
            }
            if (ok) lf_filter_sample<FL, LUMA, PASS == 0 ? 1 : 0, PASS == 0 ? 0 : 1>(R, x, y, S.str[lvl], bd);
        }
    }
}

template <int PASS, typename L, bool EDGE>
__device__ __forceinline__ void lf_pass(const LfRegion<L, EDGE> &R, int32_t X0, const LfShared &S, bool luma,
                                        uint32_t bd) {
    const uint16_t *l0 = PASS == 0 ? S.list0[0] : S.list1[0];
    const uint16_t *l1 = PASS == 0 ? S.list0[1] : S.list1[1];
    const uint16_t *l2 = PASS == 0 ? S.list0[2] : S.list1[2];
    const uint32_t n0 = __builtin_amdgcn_readfirstlane(S.cnt[PASS][0]);
    const uint32_t n1 = __builtin_amdgcn_readfirstlane(S.cnt[PASS][1]);
    const uint32_t n2 = __builtin_amdgcn_readfirstlane(S.cnt[PASS][2]);
    if (luma) {
        lf_bucket<PASS, 2, true>(R, l0, n0, X0, S, bd);
        lf_bucket<PASS, 3, true>(R, l1, n1, X0, S, bd);
        lf_bucket<PASS, 4, true>(R, l2, n2, X0, S, bd);
    } else {
        lf_bucket<PASS, 2, false>(R, l0, n0, X0, S, bd);
        lf_bucket<PASS, 3, false>(R, l1, n1, X0, S, bd);
    }
}

// Four consecutive samples of T as one vector access.
typedef uint32_t lf_u32x2 __attribute__((ext_vector_type(2)));
typedef int32_t lf_i32x4 __attribute__((ext_vector_type(4)));
template <typename T>
struct LfVec;
template <>
struct LfVec<uint8_t> {
    using type = uint32_t;
};
template <>
struct LfVec<uint16_t> {
    using type = lf_u32x2;
};
template <>
struct LfVec<int32_t> {
    using type = lf_i32x4;
};

// four samples from a vector into LDS at sm (4-byte aligned)
template <typename T, typename L>
__device__ __forceinline__ void lf_put4(L *sm, typename LfVec<T>::type v) {
    if constexpr (std::is_same<T, uint8_t>::value) {
        uint32_t *d = (uint32_t *)sm;
        d[0] = (v & 0xffu) | ((v << 8) & 0xff0000u);
        d[1] = ((v >> 16) & 0xffu) | ((v >> 8) & 0xff0000u);
    } else if constexpr (std::is_same<T, uint16_t>::value) {
        uint32_t *d = (uint32_t *)sm;
        d[0] = v.x;
        d[1] = v.y;
    } else {
        sm[0] = v.x;
        sm[1] = v.y;
        sm[2] = v.z;
        sm[3] = v.w;
    }
}
template <typename T, typename L>
__device__ __forceinline__ typename LfVec<T>::type lf_get4(const L *sm) {
    if constexpr (std::is_same<T, uint8_t>::value) {
        const uint32_t *s = (const uint32_t *)sm;
        const uint32_t a = s[0], b = s[1];
        return (a & 0xffu) | ((a >> 8) & 0xff00u) | ((b & 0xffu) << 16) | ((b << 8) & 0xff000000u);
    } else if constexpr (std::is_same<T, uint16_t>::value) {
        const uint32_t *s = (const uint32_t *)sm;
        return lf_u32x2{s[0], s[1]};
    } else {
        return lf_i32x4{sm[0], sm[1], sm[2], sm[3]};
    }
}

// One 64x64 tile of one plane: src (T) -> dst (T).
template <typename T>
__device__ __forceinline__ void lf_tile(const FrameJob &J, uint32_t plane, uint32_t ty, uint32_t tx,
                                        const G<T> *__restrict__ src, uint32_t spitch, G<T> *__restrict__ dst,
                                        uint32_t dpitch, typename LdsOf<T>::type *sm, LfShared &S) {
    using L = typename LdsOf<T>::type;
    using V = typename LfVec<T>::type;
    const int32_t rows = (int32_t)wave_u32(J.pad_rows[plane]), cols = (int32_t)wave_u32(J.pad_cols[plane]);
    const int32_t Y0 = (int32_t)ty * kLfTile, X0 = (int32_t)tx * kLfTile;
    const int32_t ry0 = Y0 - kLfHalo, rx0 = X0 - kLfHalo;
    const bool interior = ry0 >= 0 && rx0 >= 0 && ry0 + kLfRegion <= rows && rx0 + kLfRegion <= cols;
    const int tid = threadIdx.x, nt = blockDim.x;
    const uint32_t bd = wave_u32(J.bd);
    const bool luma = plane == 0;
    const uint32_t sx = luma ? 0u : wave_u32(J.sx), sy = luma ? 0u : wave_u32(J.sy);

    // 0. per-tile tables and counters
    if (tid < 32) {
        // planeSize = get_plane_residual_size(MiSize, plane).unwrap_or(MiSize)
        // (the r387 fallback keeps MiSize when the subsampled size is invalid)
        uint32_t ps = (uint32_t)tid;
        if (ps < 22) {
            const uint32_t raw = c_subsampled[ps * 4 + (sx != 0) * 2 + (sy != 0)];
            if (raw != OXAV1_BLOCK_INVALID) ps = raw;
        }
        const uint32_t bw = ps < 22 ? floor_log2(c_num4x4_w[ps] * 4u) : 2u;
        const uint32_t bh = ps < 22 ? floor_log2(c_num4x4_h[ps] * 4u) : 2u;
        S.dims[tid] = (uint8_t)(bw | bh << 4);
    } else if (tid < 48) {
        const uint32_t pass = (uint32_t)(tid - 32) >> 3, seg = (uint32_t)tid & 7u;
        const uint32_t i = luma ? pass : plane + 1;
        S.segd[pass][seg] = J.seg_data[seg][1 + i];
    } else if (tid < 54) {
        (&S.cnt[0][0])[tid - 48] = 0;
    } else if (tid >= 64 && tid < 128) {
        // strength (loop_filter.rs:502-523) of level tid - 64, scaled to the bit depth
        const int32_t lvl = tid - 64;
        const int32_t sharp = J.lf_sharp;
        const int32_t shift = sharp > 4 ? 2 : (sharp > 0 ? 1 : 0);
        const int32_t limit = sharp > 0 ? clip3(1, 9 - sharp, lvl >> shift) : max(lvl >> shift, 1);
        const int32_t blimit = 2 * (lvl + 2) + limit, thresh = lvl >> 4;
        const uint32_t bds = bd - 8;
        S.str[lvl] = (uint32_t)(limit << bds) | (uint32_t)(blimit << bds) << 12 | (uint32_t)(thresh << bds) << 24;
    }

    // 1. load the region: rows clamped to the plane (rows outside it are
    //    never read: reads clamp to the plane), four-sample groups inside the
    //    plane as one vector load, groups across the right edge per sample
    const bool vec_in = (spitch & 3u) == 0 && ((uintptr_t)src & (4 * sizeof(T) - 1)) == 0;
    if (vec_in) {
        for (int li = tid; li < kLfGroups * kLfLoadRows; li += nt) {
            const int g = li % kLfGroups, r0 = li / kLfGroups;
            const int32_t x = rx0 + 4 * g;
            const bool full = x >= 0 && x + 4 <= cols;
            const bool part = x >= 0 && x < cols && !full;
            V v[kLfRegion / kLfLoadRows];
#pragma unroll
            for (int k = 0; k < kLfRegion / kLfLoadRows; k++) {
                const int32_t y = clip3(0, rows - 1, ry0 + r0 + kLfLoadRows * k);
                if (full) v[k] = *(const G<V> *)(src + (size_t)y * spitch + x);
            }
#pragma unroll
            for (int k = 0; k < kLfRegion / kLfLoadRows; k++) {
                L *d = sm + (r0 + kLfLoadRows * k) * kLfStride + 4 * g;
                if (full) lf_put4<T>(d, v[k]);
                if (part) {
                    const int32_t y = clip3(0, rows - 1, ry0 + r0 + kLfLoadRows * k);
                    for (int j = 0; j < cols - x; j++) d[j] = (L)src[(size_t)y * spitch + x + j];
                }
            }
        }
    } else {
        // unaligned pitch: one sample per load, eight loads in flight
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
                if (i < kN) sm[ry * kLfStride + rx] = v[k];
            }
        }
    }
    __syncthreads();

    // 2. edge decisions, pushed onto the bucketed lists.  Cell (cr, e) is the
    //    plane's 4x4 cell (ry0/4 + cr, (X0-4)/4 + e); pass 0 decides its left
    //    edge for every region row, pass 1 its top edge for the tile columns
    //    and the edge rows Y0-4 .. Y0+68.  Edges whose writes cannot reach the
    //    tile are dropped: pass-0 columns X0-4 and X0+68 and pass-1 rows Y0-4
    //    and Y0+68 reach it only with filter size 16 (writes +-6).
    if (wave_u32(J.apply_lf) && wave_u32(J.lf_plane_on[plane])) {
        LfUni U;
        U.mi = (const G<uint32_t> *)wave_ptr(J.lf_mi);
        U.tx = (const G<uint8_t> *)wave_ptr(J.tx[plane]);
        U.dlf = (const G<int8_t> *)wave_ptr(J.delta_lf);
        U.sx = sx;
        U.sy = sy;
        U.mi_rows = wave_u32(J.mi_rows);
        U.mi_cols = wave_u32(J.mi_cols);
        U.fw = wave_u32(J.fw);
        U.fh = wave_u32(J.fh);
        U.tcols = wave_u32(J.tx_cols[plane]);
        U.cap = luma ? 4u : 3u;
        const uint32_t multi = wave_u32(J.lf_delta_multi);
#pragma unroll
        for (int p = 0; p < 2; p++) {
            const uint32_t i = luma ? (uint32_t)p : plane + 1;
            U.base[p] = wave_u32(J.lf_level[i]);
            U.dli[p] = multi ? i : 0u;
            uint32_t sm8 = 0;
            for (int s = 0; s < 8; s++) sm8 |= (J.seg_active[s][1 + i] ? 1u : 0u) << s;
            U.segm[p] = wave_u32(sm8);
        }
        uint64_t refd = 0;
        for (int r = 0; r < 8; r++) refd |= (uint64_t)(uint8_t)J.lf_ref_deltas[r] << (8 * r);
        U.refd = ((uint64_t)wave_u32((uint32_t)(refd >> 32)) << 32) | wave_u32((uint32_t)refd);
        U.md0 = (int32_t)wave_u32((uint32_t)(int32_t)J.lf_mode_deltas[0]);
        U.md1 = (int32_t)wave_u32((uint32_t)(int32_t)J.lf_mode_deltas[1]);
        U.delta_en = wave_u32(J.lf_delta_enabled);
        const int32_t by0 = ry0 >> 2, bx0 = (X0 - 4) >> 2;
        for (int base = 0; base < kLfCells; base += nt) {
            const int c = base + tid;
            const int cr = c / kLfEdges, e = c - cr * kLfEdges;
            uint32_t d0 = 0, d1 = 0;
            if (c < kLfCells) {
                const bool want1 = cr >= 2 && cr <= 20 && e >= 1 && e <= 16;
                lf_decide<true, true>(U, S, by0 + cr, bx0 + e, true, want1, d0, d1);
                if ((e == 0 || e == kLfEdges - 1) && (d0 >> 6) != 2) d0 = 0;
                if ((cr == 2 || cr == 20) && (d1 >> 6) != 2) d1 = 0;
            }
            const uint16_t e0 = lf_entry((uint32_t)e, (uint32_t)cr, d0 & 63u);
            const uint16_t e1 = lf_entry((uint32_t)e, (uint32_t)cr, d1 & 63u);
#pragma unroll
            for (uint32_t b = 0; b < 3; b++) {
                if (!luma && b == 2) break;
                lf_push(&S.cnt[0][b], S.list0[b], d0 != 0 && (d0 >> 6) == b, e0);
                lf_push(&S.cnt[1][b], S.list1[b], d1 != 0 && (d1 >> 6) == b, e1);
            }
        }
    }
    __syncthreads();

    // 3./4. pass 0 (vertical edges), then pass 1 (horizontal edges)
    if (interior) {
        const LfRegion<L, false> R{sm, ry0, rx0, rows, cols};
        lf_pass<0>(R, X0, S, luma, bd);
        __syncthreads();
        lf_pass<1>(R, X0, S, luma, bd);
    } else {
        const LfRegion<L, true> R{sm, ry0, rx0, rows, cols};
        lf_pass<0>(R, X0, S, luma, bd);
        __syncthreads();
        lf_pass<1>(R, X0, S, luma, bd);
    }
    __syncthreads();

    // 5. store the tile, four samples per lane where the pitch allows
    const bool vec_out = (dpitch & 3u) == 0 && ((uintptr_t)dst & (4 * sizeof(T) - 1)) == 0;
    if (vec_out) {
        for (int gi = tid; gi < kLfTile * kLfTile / 4; gi += nt) {
            const int g = gi & 15, yy = gi >> 4;  // 16 four-sample groups per tile row
            const int32_t x = X0 + 4 * g, y = Y0 + yy;
            const bool full = x + 4 <= cols;
            const bool part = x < cols && !full;
            if (y < rows) {
                const L *s = sm + (yy + kLfHalo) * kLfStride + kLfHalo + 4 * g;
                if (full) *(G<V> *)(dst + (size_t)y * dpitch + x) = lf_get4<T>(s);
                if (part)
                    for (int j = 0; j < cols - x; j++) dst[(size_t)y * dpitch + x + j] = (T)s[j];
            }
        }
    } else {
        for (int base = 0; base < kLfTile * kLfTile; base += nt) {
            const int i = base + tid;
            const int yy = i >> 6, xx = i & 63;
            const int32_t y = Y0 + yy, x = X0 + xx;
            if (i < kLfTile * kLfTile && y < rows && x < cols)
                dst[(size_t)y * dpitch + x] = (T)sm[(yy + kLfHalo) * kLfStride + (xx + kLfHalo)];
        }
    }
}

}  // namespace oxav1
