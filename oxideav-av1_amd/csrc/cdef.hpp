// §7.15 CDEF on gfx950: one workgroup per 64x64 filter unit (all planes).
//
// Reference: cdef.rs:180-606.  CDEF reads only the deblocked frame and
// writes a separate frame (cdef.rs:231-251), so 8x8 blocks are independent.
// The workgroup stages the unit's luma and chroma source tiles with their
// 2-sample borders in LDS, runs the §7.15.2 direction search with one lane
// per 8x8 block, then filters every sample with the block's strengths.
// Tap availability is "inside the plane buffer" (cdef.rs:501-511, 576-583);
// units whose bordered tile lies inside the plane skip the per-tap checks.
#pragma once

#include "common.hpp"

namespace oxav1 {

constexpr int kCdefUnit = 64;           // luma samples per unit side
constexpr int kCdefB = 2;               // tap reach
constexpr int kCdefLw = kCdefUnit + 2 * kCdefB;  // 68
// The engine's u16 LDS tile is stored twice, the second copy kCdefDual
// (odd) elements after the first, so a column pair starting at an odd
// element of the first copy starts at an even one of the second: every pair
// the packed filter reads is an aligned dword.
constexpr int kCdefDual = kCdefLw * kCdefLw + 1;
typedef uint32_t cdef_u32x2 __attribute__((ext_vector_type(2)));

__constant__ int8_t c_cdef_dir[8][2][2] = {
    {{-1, 1}, {-2, 2}}, {{0, 1}, {-1, 2}}, {{0, 1}, {0, 2}}, {{0, 1}, {1, 2}},
    {{1, 1}, {2, 2}},   {{1, 0}, {2, 1}},  {{1, 0}, {2, 0}}, {{1, 0}, {2, -1}}};
__constant__ uint8_t c_cdef_uv_dir[2][2][8] = {{{0, 1, 2, 3, 4, 5, 6, 7}, {1, 2, 2, 2, 3, 4, 6, 0}},
                                               {{7, 0, 2, 4, 5, 6, 6, 6}, {0, 1, 2, 3, 4, 5, 6, 7}}};
constexpr uint32_t kCdefDiv[9] = {0, 840, 420, 280, 210, 168, 140, 120, 105};  // Div_Table


// Per 8x8 block decisions shared by the unit's threads.
struct CdefBlk {
    int8_t dir_y, dir_uv;
    uint8_t active;  // 0: copy only
    uint8_t pad;
    int16_t pri_y, sec_y, pri_uv, sec_uv;
};

// Samples per partial sum of the direction search: a diagonal of the 8x8
// block holds min(k + 1, 15 - k) samples; the half-slope lines 2 * (the
// number of rows they cross).
constexpr int32_t kCdefCnt15[15] = {1, 2, 3, 4, 5, 6, 7, 8, 7, 6, 5, 4, 3, 2, 1};
constexpr int32_t kCdefCnt11[11] = {2, 4, 6, 8, 8, 8, 8, 8, 6, 4, 2};

// cdef_direction (cdef.rs:366-449), costs of two of the eight directions
// for one 8x8 block read from the LDS tile: quarter 0 -> directions 0 and 2,
// 1 -> 4 and 6, 2 -> 1 and 3, 3 -> 5 and 7 (stored at cost[d]).  Four threads
// share a block so each keeps at most 26 partial sums in registers.  The
// block rows and columns are clamped to the staged (buffer-bounded) extent
// lh x lw, as the reference clamps its reads to the plane (cdef.rs:386-391);
// CLAMP = false for blocks inside it (constant LDS offsets).  The partial
// sums add the shifted samples and take the -128 of every sample off once
// per sum (each sum's sample count is fixed).
// x * K; K24: 0 <= x < 2^24, on the 24-bit multiplier.
template <bool K24, uint32_t K>
__device__ __forceinline__ int32_t cdef_kmul(int32_t x) {
    return K24 ? (int32_t)mul_u24k<K>((uint32_t)x) : x * (int32_t)K;
}

template <typename L, int Q, bool CLAMP>
__device__ __forceinline__ void cdef_dir_quarter(const S3<L> *t, int ls, int lh, int lw, int y8, int x8, int32_t shift,
                                                 S3<int32_t> *cost) {
    const S3<L> *base = t + mul_u24(y8 + kCdefB, ls) + x8 + kCdefB;
    // With 16-bit tiles the partial sums hold at most eight samples of
    // 8 bits less 128 (|sum| <= 1024), so squares (<= 2^20), pairs of them
    // times Div_Table (< 2^31) and the eight-square sums times 105 (< 2^31)
    // run on the 24-bit multiplier; int32 tiles keep 32-bit products.
    constexpr bool k24 = sizeof(L) <= 2;
    auto sq = [](int32_t p, int32_t acc) -> int32_t { return k24 ? mad_i24(p, p, acc) : p * p + acc; };
    auto px = [&](int i, int j) -> int32_t {
        if (CLAMP) {
            const int yy = min(y8 + i, lh - 1), xx = min(x8 + j, lw - 1);
            return (int32_t)t[mul_u24(yy + kCdefB, ls) + xx + kCdefB] >> shift;
        }
        return (int32_t)base[i * ls + j] >> shift;
    };
    if constexpr (Q < 2) {
        // Q 0: d = 0 (i + j) and d = 2 (rows); Q 1: d = 4 (7 + i - j) and
        // d = 6 (columns).  Diagonals: 15 partial sums, outer pairs weighted
        // by Div_Table.
        int32_t pd[15], pa[8];
#pragma unroll
        for (int k = 0; k < 15; k++) pd[k] = -128 * kCdefCnt15[k];
#pragma unroll
        for (int k = 0; k < 8; k++) pa[k] = -128 * 8;
#pragma unroll
        for (int i = 0; i < 8; i++) {
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int32_t v = px(i, j);
                pd[Q == 0 ? i + j : 7 + i - j] += v;
                pa[Q == 0 ? i : j] += v;
            }
            // one row's samples in flight at a time (registers)
            __builtin_amdgcn_sched_barrier(0);
        }
        // Div_Table[k + 1] weighs lines k and 14 - k
        int32_t cd = cdef_kmul<k24, 105>(sq(pd[7], 0)), ca = 0;
        cd += cdef_kmul<k24, kCdefDiv[1]>(sq(pd[0], sq(pd[14], 0)));
        cd += cdef_kmul<k24, kCdefDiv[2]>(sq(pd[1], sq(pd[13], 0)));
        cd += cdef_kmul<k24, kCdefDiv[3]>(sq(pd[2], sq(pd[12], 0)));
        cd += cdef_kmul<k24, kCdefDiv[4]>(sq(pd[3], sq(pd[11], 0)));
        cd += cdef_kmul<k24, kCdefDiv[5]>(sq(pd[4], sq(pd[10], 0)));
        cd += cdef_kmul<k24, kCdefDiv[6]>(sq(pd[5], sq(pd[9], 0)));
        cd += cdef_kmul<k24, kCdefDiv[7]>(sq(pd[6], sq(pd[8], 0)));
#pragma unroll
        for (int k = 0; k < 8; k++) ca = sq(pa[k], ca);
        cost[Q == 0 ? 0 : 4] = cd;
        cost[Q == 0 ? 2 : 6] = cdef_kmul<k24, 105>(ca);
    } else {
        // Q 2: d = 1 (i + j/2) and d = 3 (3 + i - j/2); Q 3: d = 5
        // (3 - i/2 + j) and d = 7 (i/2 + j).  11 partial sums, the middle
        // five weighted 105.
        int32_t pa[11], pb[11];
#pragma unroll
        for (int k = 0; k < 11; k++) pa[k] = pb[k] = -128 * kCdefCnt11[k];
#pragma unroll
        for (int i = 0; i < 8; i++) {
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int32_t v = px(i, j);
                pa[Q == 2 ? i + j / 2 : 3 - i / 2 + j] += v;
                pb[Q == 2 ? 3 + i - j / 2 : i / 2 + j] += v;
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        auto half_cost = [&](const int32_t *p) {
            int32_t c = 0;
#pragma unroll
            for (int k = 0; k < 5; k++) c = sq(p[3 + k], c);
            c = cdef_kmul<k24, 105>(c);
            c += cdef_kmul<k24, kCdefDiv[2]>(sq(p[0], sq(p[10], 0)));
            c += cdef_kmul<k24, kCdefDiv[4]>(sq(p[1], sq(p[9], 0)));
            c += cdef_kmul<k24, kCdefDiv[6]>(sq(p[2], sq(p[8], 0)));
            return c;
        };
        cost[Q == 2 ? 1 : 5] = half_cost(pa);
        cost[Q == 2 ? 3 : 7] = half_cost(pb);
    }
}

// FULL: the tile is staged over the whole unit (rows and columns past the
// plane hold the plane's clamped samples), so no block needs the clamps.
template <typename L, bool FULL = false>
__device__ __forceinline__ void cdef_dir_costs(const S3<L> *t, int ls, int lh, int lw, int y8, int x8, int32_t shift,
                                               int quarter, S3<int32_t> *cost) {
    const bool in = FULL || (y8 + 8 <= lh && x8 + 8 <= lw);
    switch (quarter) {  // uniform per wave
    case 0:
        if (in) cdef_dir_quarter<L, 0, false>(t, ls, lh, lw, y8, x8, shift, cost);
        else if constexpr (!FULL) cdef_dir_quarter<L, 0, true>(t, ls, lh, lw, y8, x8, shift, cost);
        break;
    case 1:
        if (in) cdef_dir_quarter<L, 1, false>(t, ls, lh, lw, y8, x8, shift, cost);
        else if constexpr (!FULL) cdef_dir_quarter<L, 1, true>(t, ls, lh, lw, y8, x8, shift, cost);
        break;
    case 2:
        if (in) cdef_dir_quarter<L, 2, false>(t, ls, lh, lw, y8, x8, shift, cost);
        else if constexpr (!FULL) cdef_dir_quarter<L, 2, true>(t, ls, lh, lw, y8, x8, shift, cost);
        break;
    default:
        if (in) cdef_dir_quarter<L, 3, false>(t, ls, lh, lw, y8, x8, shift, cost);
        else if constexpr (!FULL) cdef_dir_quarter<L, 3, true>(t, ls, lh, lw, y8, x8, shift, cost);
        break;
    }
}

// yDir = first argmax (ties to the lowest direction), var (cdef.rs:434-447)
__device__ __forceinline__ void cdef_pick_dir(const int32_t *cost, int32_t *dir_out, int32_t *var_out) {
    int32_t best = 0, dir = 0;
#pragma unroll
    for (int k = 0; k < 8; k++)
        if (cost[k] > best) {
            best = cost[k];
            dir = k;
        }
    *dir_out = dir;
    *var_out = (best - cost[(dir + 4) & 7]) >> 10;
}

// constrain (cdef.rs:592-606) with the per-block damping adjustment
// precomputed (adj capped at 31, which shifts any |diff| < 2^31 to 0).
__device__ __forceinline__ int32_t cdef_constrain(int32_t diff, int32_t thr, int32_t adj) {
    const int32_t ad = abs(diff);
    const int32_t v = min(ad, max(thr - (ad >> adj), 0));
    return diff < 0 ? -v : v;
}
__device__ __forceinline__ int32_t cdef_adj(int32_t damping, int32_t thr) {
    if (thr == 0) return 0;
    return min(max(0, damping - (int32_t)floor_log2((uint32_t)thr)), 31);
}

// Taps of one 8x8 block for the plane being filtered, resolved once per
// block: LDS offsets of the six "+" taps (primary k=0,1; secondary dir+2 and
// dir-2 at k=0,1; the "-" taps are the negated offsets), their (dy, dx) for
// the per-tap availability test, strengths and damping shifts.
struct CdefTaps {
    int16_t off[6];
    int16_t pri, sec;
    uint8_t pri_adj, sec_adj, pt0, pt1;
    uint8_t active, pad[3];
    int8_t dy[6], dx[6];
};

__device__ __forceinline__ CdefTaps cdef_taps(const CdefBlk &b, bool chroma, int32_t damping,
                                             int32_t coeff_shift, int ls) {
    CdefTaps t{};
    const int32_t pri = chroma ? b.pri_uv : b.pri_y;
    const int32_t sec = chroma ? b.sec_uv : b.sec_y;
    const int32_t dir = chroma ? b.dir_uv : b.dir_y;
    t.active = b.active;
    t.pri = (int16_t)pri;
    t.sec = (int16_t)sec;
    t.pri_adj = (uint8_t)cdef_adj(damping, pri);
    t.sec_adj = (uint8_t)cdef_adj(damping, sec);
    const int pri_row = (pri >> coeff_shift) & 1;  // Cdef_Pri_Taps[(priStr >> coeffShift) & 1]
    t.pt0 = pri_row ? 3 : 4;
    t.pt1 = pri_row ? 3 : 2;
    const int32_t ds[6] = {dir, dir, (dir + 2) & 7, (dir - 2) & 7, (dir + 2) & 7, (dir - 2) & 7};
    const int ks[6] = {0, 1, 0, 0, 1, 1};
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const int32_t dy = c_cdef_dir[ds[i]][ks[i]][0], dx = c_cdef_dir[ds[i]][ks[i]][1];
        t.dy[i] = (int8_t)dy;
        t.dx[i] = (int8_t)dx;
        t.off[i] = (int16_t)(dy * ls + dx);
    }
    return t;
}

// Filter one plane of a unit from an LDS tile with a kCdefB border
// (cdef_filter_block, cdef.rs:465-590).  Lanes own columns: 1 << row_log2
// consecutive threads cover one tile row, and each thread walks down its
// column in steps of blockDim.x >> row_log2 rows.  Every filtered sample
// goes to sink(yy, xx, py, px, value) (tile and plane coordinates).
// CHECK: per-tap "inside the plane" test (units touching the frame edge).
template <typename L, bool CHECK, typename Sink>
__device__ __forceinline__ void cdef_filter_tile(const L *t, int ls, int tw, int th, int32_t ox, int32_t oy, int32_t pw,
                                                 int32_t ph, int32_t bw_log2, int32_t bh_log2, const CdefTaps *taps,
                                                 int32_t row_log2, Sink &&sink) {
    const int xx = threadIdx.x & ((1 << row_log2) - 1);
    const int ystep = blockDim.x >> row_log2;
    if (xx >= tw) return;
    const int32_t px = ox + xx;
    if (px >= pw) return;
    // first row: the wave's (scalar) base plus the lane's sub-row when a
    // wave covers several rows
    const int y_first = (__builtin_amdgcn_readfirstlane(threadIdx.x >> 6) << (6 - row_log2)) +
                        ((int)(threadIdx.x & 63) >> row_log2);
    for (int yy = y_first; yy < th; yy += ystep) {
        const int32_t py = oy + yy;
        if (py >= ph) break;
        const CdefTaps b = taps[(yy >> bh_log2) * 8 + (xx >> bw_log2)];
        const L *c = t + mul_u24(yy + kCdefB, ls) + (xx + kCdefB);
        const int32_t x = (int32_t)c[0];
        int32_t out = x;
        if (b.active) {
            int32_t sum = 0, mx = x, mn = x;
#pragma unroll
            for (int i = 0; i < 6; i++) {
                const uint32_t thr = (uint32_t)(i < 2 ? b.pri : b.sec);
                const uint32_t adj = i < 2 ? b.pri_adj : b.sec_adj;
                const int32_t w = i == 0 ? b.pt0 : (i == 1 ? b.pt1 : (i < 4 ? 2 : 1));
#pragma unroll
                for (int sg = 0; sg < 2; sg++) {
                    if (CHECK) {
                        const int32_t dy = sg ? -b.dy[i] : b.dy[i], dx = sg ? -b.dx[i] : b.dx[i];
                        if (!(py + dy >= 0 && py + dy < ph && px + dx >= 0 && px + dx < pw)) continue;
                    }
                    const int32_t v = (int32_t)c[sg ? -b.off[i] : b.off[i]];
                    // constrain (cdef.rs:592-606): sign(d) * min(|d|, t) with
                    // t = max(0, thr - (|d| >> adj)) is d clamped to [-t, t]
                    const int32_t d = v - x;
                    uint32_t ad;
                    if constexpr (std::is_same<L, uint16_t>::value)
                        ad = __builtin_amdgcn_sad_u16((uint32_t)v, (uint32_t)x, 0u);  // samples < 2^16
                    else
                        ad = (uint32_t)abs(d);
                    const int32_t tt = (int32_t)__builtin_elementwise_sub_sat(thr, ad >> adj);
                    // d clamped to [-tt, tt], tt >= 0; |d| < 2^16 for 16-bit tiles
                    const int32_t dc = med3_i32(d, -tt, tt);
                    sum = sizeof(L) <= 2 ? mad_i24(w, dc, sum) : sum + w * dc;
                    mx = max(mx, v);
                    mn = min(mn, v);
                }
            }
            const int32_t delta = (8 + sum - (sum < 0 ? 1 : 0)) >> 4;
            out = clip3(mn, mx, x + delta);
        }
        sink(yy, xx, py, px, out);
    }
}

// Two samples per lane: the interior path (every tap inside the plane) of
// the engine's u8/u16 frames, with the LDS tile in u16.  Lanes own column
// pairs (xx, xx + 1), xx even, which always lie in one block (blocks are at
// least 4 samples wide), so both samples share the block's taps, strengths
// and damping shifts.  Every intermediate fits a signed 16-bit half: tap
// differences |d| < 2^12, thresholds <= 240 (15 << 4 at 12-bit), and the
// weighted sum |sum| <= 12 * 240 + 12 * 64 = 3648; the damping shift is
// capped at 15, which like the scalar cap of 31 leaves 0 for any |d| <
// 2^15.  The pair of a tap at an odd offset is an unaligned dword of LDS.
// Per tap, constrain (cdef.rs:592-606) is d clamped to [-t, t] with t =
// max(0, thr - (|d| >> adj)) (an unsigned saturating subtract).
typedef short cdef_s16x2 __attribute__((ext_vector_type(2)));
typedef unsigned short cdef_u16x2 __attribute__((ext_vector_type(2)));

// A block's taps for the packed path: byte offsets of its twelve taps (+
// and - of each CdefTaps::off) from a pair's own dword in the first copy
// (into the first copy for even element offsets, into the second for odd
// ones), the strengths and damping shifts (capped at 15) and the primary
// weights; bit 7 of pt0 says the block is filtered.  32 bytes, so a block's
// entry is found with a shift.
struct CdefTapsPk {
    int16_t boff[12];
    uint16_t pri, sec;
    uint8_t pri_adj, sec_adj, pt0, pt1;
};
static_assert(sizeof(CdefTapsPk) == 32, "packed tap entries are 32 bytes");

__device__ __forceinline__ CdefTapsPk cdef_taps_pk(const CdefTaps &t) {
    CdefTapsPk k{};
#pragma unroll
    for (int i = 0; i < 6; i++)
#pragma unroll
        for (int sg = 0; sg < 2; sg++) {
            const int32_t o = sg ? -t.off[i] : t.off[i];
            k.boff[2 * i + sg] = (int16_t)(2 * o + ((o & 1) ? 2 * kCdefDual : 0));
        }
    k.pri = (uint16_t)t.pri;
    k.sec = (uint16_t)t.sec;
    k.pri_adj = (uint8_t)min((int32_t)t.pri_adj, 15);
    k.sec_adj = (uint8_t)min((int32_t)t.sec_adj, 15);
    k.pt0 = (uint8_t)(t.pt0 | (t.active ? 0x80u : 0u));
    k.pt1 = t.pt1;
    return k;
}

// The unit's per-block tables share one LDS region: the direction costs
// (64 x 8 words) until the blocks' decisions are made, then either the
// scalar or the packed taps of the plane being filtered.
union CdefTables {
    int32_t costs[64 * 8];
    CdefTaps taps[64];
    CdefTapsPk pk[64];
};

// PRI / SEC: some block of the plane has a nonzero primary / secondary
// strength (both are per unit, up to the per-block primary adjustment to 0).
// A tap whose strength is 0 adds 0 (constrain returns 0), and with only
// primary or only secondary taps the weights sum to 12 < 16, so
// x + ((8 + sum - (sum < 0)) >> 4) stays between the smallest and the
// largest tap (x included) and the final clip is the identity: the
// instances without one kind skip its taps and the min/max tracking.
template <typename T, bool PRI = true, bool SEC = true>
__device__ __forceinline__ void cdef_filter_plane_pk(const uint16_t *t, int ls, int tw, int th, int32_t ox, int32_t oy,
                                                     int32_t bw_log2, int32_t bh_log2, const CdefTapsPk *taps,
                                                     int32_t row_log2, G<T> *__restrict__ dst, uint32_t dpitch) {
    using s16x2 = cdef_s16x2;
    using u16x2 = cdef_u16x2;
    const int pr_log2 = row_log2 - 1;  // column pairs per tile row (log2)
    const int xx = 2 * ((int)threadIdx.x & ((1 << pr_log2) - 1));
    const int ystep = blockDim.x >> pr_log2;
    if (xx >= tw) return;
    const int y_first = (__builtin_amdgcn_readfirstlane(threadIdx.x >> 6) << (6 - pr_log2)) +
                        ((int)(threadIdx.x & 63) >> pr_log2);
    auto pair_at = [](const uint16_t *q, int32_t byte_off) -> s16x2 {
        return __builtin_bit_cast(s16x2, *(const uint32_t *)((const char *)q + byte_off));
    };
    auto splat = [](int32_t v) -> uint32_t { return (uint32_t)v | (uint32_t)v << 16; };  // 0 <= v < 2^16
    for (int yy = y_first; yy < th; yy += ystep) {
        const CdefTapsPk b = taps[(yy >> bh_log2) * 8 + (xx >> bw_log2)];
        const uint16_t *c = t + mul_u24(yy + kCdefB, ls) + (xx + kCdefB);
        const s16x2 x = pair_at(c, 0);
        s16x2 res = x;
        if (b.pt0 & 0x80u) {
            const u16x2 thr_p = __builtin_bit_cast(u16x2, splat(b.pri));
            const u16x2 thr_s = __builtin_bit_cast(u16x2, splat(b.sec));
            const u16x2 adj_p = __builtin_bit_cast(u16x2, splat(b.pri_adj));
            const u16x2 adj_s = __builtin_bit_cast(u16x2, splat(b.sec_adj));
            s16x2 acc[3] = {{0, 0}, {0, 0}, {0, 0}};  // primary k = 0, primary k = 1, secondary
            s16x2 sec1 = {0, 0};                      // secondary k = 1 (weight 1; k = 0 weighs 2)
            s16x2 mx = x, mn = x;
#pragma unroll
            for (int i = 0; i < 6; i++) {
                if ((i < 2 && !PRI) || (i >= 2 && !SEC)) continue;
                const u16x2 thr = i < 2 ? thr_p : thr_s;
                const u16x2 adj = i < 2 ? adj_p : adj_s;
#pragma unroll
                for (int sg = 0; sg < 2; sg++) {
                    const s16x2 v = pair_at(c, b.boff[2 * i + sg]);
                    const s16x2 d = v - x;
                    const s16x2 ad = __builtin_elementwise_max(d, x - v);
                    const u16x2 tt = __builtin_elementwise_sub_sat(thr, __builtin_bit_cast(u16x2, ad) >> adj);
                    const s16x2 ts = __builtin_bit_cast(s16x2, tt);
                    const s16x2 cv = __builtin_elementwise_max(__builtin_elementwise_min(d, ts), -ts);
                    if (i < 2)
                        acc[i] += cv;
                    else if (i < 4)
                        acc[2] += cv;
                    else
                        sec1 += cv;
                    if (PRI && SEC) {
                        mx = __builtin_elementwise_max(mx, v);
                        mn = __builtin_elementwise_min(mn, v);
                    }
                }
            }
            const s16x2 w0 = __builtin_bit_cast(s16x2, splat(b.pt0 & 0x7fu));
            const s16x2 w1 = __builtin_bit_cast(s16x2, splat(b.pt1));
            const s16x2 sum = acc[0] * w0 + acc[1] * w1 + (acc[2] + acc[2]) + sec1;
            // (8 + sum - (sum < 0)) >> 4, arithmetic shifts on each half
            const s16x2 neg = sum >> (s16x2){15, 15};
            const s16x2 delta = (sum + neg + (s16x2){8, 8}) >> (s16x2){4, 4};
            res = (PRI && SEC) ? __builtin_elementwise_min(__builtin_elementwise_max(x + delta, mn), mx) : x + delta;
        }
        const int32_t py = oy + yy, px = ox + xx;
        const uint32_t r = __builtin_bit_cast(uint32_t, res);
        if constexpr (std::is_same<T, uint16_t>::value) {
            *(G<uint32_t> *)(dst + (size_t)py * dpitch + px) = r;  // px and the pitch are even
        } else {
            *(G<uint16_t> *)(dst + (size_t)py * dpitch + px) = (uint16_t)((r & 0xffu) | ((r >> 8) & 0xff00u));
        }
    }
}

template <typename T>
__device__ __forceinline__ void cdef_filter_plane_pk_s(bool pri, bool sec, const uint16_t *t, int ls, int tw, int th,
                                                       int32_t ox, int32_t oy, int32_t bw_log2, int32_t bh_log2,
                                                       const CdefTapsPk *taps, int32_t row_log2, G<T> *__restrict__ dst,
                                                       uint32_t dpitch) {
    if (pri && sec)
        cdef_filter_plane_pk<T, true, true>(t, ls, tw, th, ox, oy, bw_log2, bh_log2, taps, row_log2, dst, dpitch);
    else if (pri)
        cdef_filter_plane_pk<T, true, false>(t, ls, tw, th, ox, oy, bw_log2, bh_log2, taps, row_log2, dst, dpitch);
    else
        cdef_filter_plane_pk<T, false, true>(t, ls, tw, th, ox, oy, bw_log2, bh_log2, taps, row_log2, dst, dpitch);
}

template <typename L, typename T, bool CHECK>
__device__ __forceinline__ void cdef_filter_plane(const L *t, int ls, int tw, int th, int32_t ox, int32_t oy, int32_t pw,
                                  int32_t ph, int32_t bw_log2, int32_t bh_log2, const CdefTaps *taps,
                                  int32_t row_log2, G<T> *__restrict__ dst, uint32_t dpitch) {
    cdef_filter_tile<L, CHECK>(t, ls, tw, th, ox, oy, pw, ph, bw_log2, bh_log2, taps, row_log2,
                               [&](int, int, int32_t py, int32_t px, int32_t v) { dst[(size_t)py * dpitch + px] = (T)v; });
}

// Exact i / d for 0 <= i < 2^20 / d with m = ceil(2^20 / d), d <= 128.
__device__ __forceinline__ int div_magic(int i, uint32_t m) {
    return (int)(mul_u24((uint32_t)i, m) >> 20);
}

// Stage a tile with its kCdefB border: t[yy * ls + xx] = src at (oy + yy -
// 2, ox + xx - 2), both coordinates clamped to the plane (out-of-plane
// entries are never used as taps; the clamp keeps every load in bounds).
// DUAL: the u16 tile is written a second time kCdefDual elements on.  Where
// the pitch allows, four plane columns (starting at a multiple of four) come
// from one vector load and land with aligned dword stores.
template <typename L, typename T, bool DUAL = false>
__device__ __forceinline__ void cdef_load(L *t, int tw, int th, int32_t ox, int32_t oy, int32_t pw, int32_t ph,
                          const G<T> *__restrict__ src, uint32_t spitch) {
    const int ls = tw + 2 * kCdefB, lh = th + 2 * kCdefB;
    const bool vec = (spitch & 3u) == 0 && ((uintptr_t)src & (4 * sizeof(T) - 1)) == 0 && (ox & 3) == 0;
    if (!vec) {
        const uint32_t m = ((1u << 20) + (uint32_t)ls - 1) / (uint32_t)ls;
        const int n = ls * lh;
#pragma unroll 4
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            const int yy = div_magic(i, m), xx = i - (int)mul_u24(yy, ls);
            const int32_t y = oy + yy - kCdefB, x = ox + xx - kCdefB;
            const int32_t yc = clip3(0, ph - 1, y), xc = clip3(0, pw - 1, x);
            const L v = (L)src[(size_t)yc * spitch + xc];
            t[i] = v;
            if (DUAL) t[kCdefDual + i] = v;
        }
        return;
    }
    // group g covers plane columns ox - 4 + 4g .. +3, tile columns 4g - 2 .. 4g + 1
    const int ng = (ls + 5) >> 2;
    const uint32_t m = ((1u << 20) + (uint32_t)ng - 1) / (uint32_t)ng;
    const int n = ng * lh;
    if constexpr (sizeof(L) == 2 && sizeof(T) <= 2) {
        // u8/u16 samples: thread t loads group column t % ng of rows t / ng,
        // t / ng + rpp, ... (rpp = rows per pass), up to five rows at once
        // (a whole 68-row unit), one aligned vector load per group, always
        // inside its row (a group past either plane edge loads the nearest
        // group inside, ga, and picks its clamped samples from it at the
        // store); the samples stay packed in registers until the LDS stores
        constexpr int kR = 5;
        using W = typename std::conditional<sizeof(T) == 1, uint32_t, cdef_u32x2>::type;
        const int tid = (int)threadIdx.x;
        const int rpp = div_magic((int)blockDim.x, m);  // m = ceil(2^20 / ng)
        const int r0 = div_magic(tid, m), g = tid - (int)mul_u24(r0, ng);
        if (r0 >= rpp) return;
        const int32_t gx = ox - 4 + 4 * g;
        const int32_t ga = clip3(0, (pw - 1) & ~3, gx);  // the last group holding a plane column at most
        const int x0 = 4 * g - kCdefB;
        // sample k is plane column clip3(0, pw - 1, gx + k), at position
        // j_k = that - ga (0..3) of the loaded group
        const bool edge = gx < 0 || gx + 3 >= pw;
        for (int yb = r0; yb < lh; yb += kR * rpp) {
            W w[kR];
#pragma unroll
            for (int k = 0; k < kR; k++) {
                const int yy = yb + k * rpp;
                if (yy < lh) w[k] = *(const G<W> *)(src + (size_t)clip3(0, ph - 1, oy + yy - kCdefB) * spitch + ga);
            }
#pragma unroll
            for (int k = 0; k < kR; k++) {
                const int yy = yb + k * rpp;
                if (yy >= lh) continue;
                const int e = (int)mul_u24(yy, ls) + x0;
                // samples 0..3 of the group as (0, 1) and (2, 3) pairs
                uint32_t lo, hi;
                if constexpr (sizeof(T) == 1) {
                    lo = (w[k] & 0xffu) | ((w[k] << 8) & 0xff0000u);
                    hi = ((w[k] >> 16) & 0xffu) | ((w[k] >> 8) & 0xff0000u);
                } else {
                    lo = w[k].x;
                    hi = w[k].y;
                }
                if (edge) {
                    uint32_t s4[4];
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const int32_t j = clip3(0, pw - 1, gx + q) - ga;
                        s4[q] = ((j < 2 ? lo : hi) >> (16 * (j & 1))) & 0xffffu;
                    }
                    lo = s4[0] | s4[1] << 16;
                    hi = s4[2] | s4[3] << 16;
                }
                if (x0 >= 0 && x0 + 3 < ls && !(e & 1)) {
                    uint32_t *d = (uint32_t *)(t + e);
                    d[0] = lo;
                    d[1] = hi;
                    if (DUAL) {  // kCdefDual is odd: e + kCdefDual + 1 is even
                        L *q = t + kCdefDual + e;
                        q[0] = (L)lo;
                        *(uint32_t *)(q + 1) = __builtin_amdgcn_alignbit(hi, lo, 16);
                        q[3] = (L)(hi >> 16);
                    }
                    continue;
                }
                const uint32_t s4[4] = {lo & 0xffffu, lo >> 16, hi & 0xffffu, hi >> 16};
#pragma unroll
                for (int q = 0; q < 4; q++)
                    if (x0 + q >= 0 && x0 + q < ls) {
                        t[e + q] = (L)s4[q];
                        if (DUAL) t[kCdefDual + e + q] = (L)s4[q];
                    }
            }
        }
        return;
    }
    constexpr int kR = 2;  // groups per thread and round (more raise the kernel's register use)
    for (int i0 = 0; i0 < n; i0 += kR * (int)blockDim.x) {
        int32_t v[kR][4];
        int yr[kR], gr[kR];
#pragma unroll
        for (int r = 0; r < kR; r++) {
            const int i = i0 + r * (int)blockDim.x + (int)threadIdx.x;
            yr[r] = -1;
            if (i < n) {
                const int yy = div_magic(i, m), g = i - (int)mul_u24(yy, ng);
                yr[r] = yy;
                gr[r] = g;
                const G<T> *row = src + (size_t)clip3(0, ph - 1, oy + yy - kCdefB) * spitch;
                const int32_t gx = ox - 4 + 4 * g;
                if (gx >= 0 && gx + 3 < pw) {
                    typedef int32_t i32x4 __attribute__((ext_vector_type(4)));
                    const i32x4 w = *(const G<i32x4> *)(row + gx);
                    v[r][0] = w.x;
                    v[r][1] = w.y;
                    v[r][2] = w.z;
                    v[r][3] = w.w;
                } else {
#pragma unroll
                    for (int k = 0; k < 4; k++) v[r][k] = (int32_t)row[clip3(0, pw - 1, gx + k)];
                }
            }
        }
#pragma unroll
        for (int r = 0; r < kR; r++) {
            if (yr[r] < 0) continue;
            const int x0 = 4 * gr[r] - kCdefB, e = (int)mul_u24(yr[r], ls) + x0;
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (x0 + k >= 0 && x0 + k < ls) {
                    t[e + k] = (L)v[r][k];
                    if (DUAL) t[kCdefDual + e + k] = (L)v[r][k];
                }
        }
    }
}

// The engine's u8/u16 frames filter interior units two samples per lane.
template <typename T>
constexpr bool kCdefPacked = std::is_same<T, uint8_t>::value || std::is_same<T, uint16_t>::value;

// One 64x64 unit (ur, uc in 64-luma units): deb -> cdf for all planes.
template <typename T>
__device__ __forceinline__ void cdef_unit(const FrameJob &J, uint32_t ur, uint32_t uc, typename LdsOf<T>::type *lds,
                          CdefBlk *blk, CdefTables *tabs) {
    int32_t *costs = tabs->costs;
    CdefTaps *taps = tabs->taps;
    using L = typename LdsOf<T>::type;
    const int tid = threadIdx.x;
    const uint32_t r0 = ur * 16, c0 = uc * 16;  // mi coordinates
    const int32_t coeff_shift = (int32_t)J.bd - 8;
    const int8_t idx = gread(J.cdef_idx, (size_t)ur * J.cdef_ucols + uc);
    const bool do_cdef = J.apply_cdef != 0;
    // luma tile
    // The 8x8 walk writes min(buffer, walked extent) (cdef.rs:185-200, 320-352);
    // taps and the direction search see the whole buffer.
    const int32_t ph0 = (int32_t)J.pad_rows[0], pw0 = (int32_t)J.pad_cols[0];
    const int32_t eh0 = (int32_t)J.cdef_ext_rows[0], ew0 = (int32_t)J.cdef_ext_cols[0];
    const int32_t oy0 = (int32_t)r0 * 4, ox0 = (int32_t)c0 * 4;
    const int tw0 = min(kCdefUnit, ew0 - ox0), th0 = min(kCdefUnit, eh0 - oy0);
    if (tw0 <= 0 || th0 <= 0) return;
    // staged extent: the whole unit, coordinates clamped to the buffer, so
    // the direction search of blocks that reach past the buffer reads the
    // clamped samples (min(y, lh - 1), min(x, lw - 1) with lh, lw the unit's
    // extent inside the buffer) without clamps of its own
    const int lw0 = kCdefUnit, lh0 = kCdefUnit;
    const int ls0 = lw0 + 2 * kCdefB;
    L *t0 = lds;
    cdef_load<L, T, kCdefPacked<T>>(t0, lw0, lh0, ox0, oy0, pw0, ph0, plane_ptr<T>(J.deb[0]), J.deb_pitch[0]);
    __syncthreads();
    // per 8x8 block: skip test (cdef.rs:266-272, out-of-grid cells count as
    // skip) and the direction search, split over the four waves (two
    // directions each) that meet through `costs`
    const int nbx = (tw0 + 7) >> 3, nby = (th0 + 7) >> 3;
    auto block_active = [&](int b) -> bool {
        const int by = b >> 3, bx = b & 7;
        if (!(by < nby && bx < nbx && do_cdef && idx >= 0)) return false;
        const uint32_t r = r0 + 2 * by, c = c0 + 2 * bx;
        if (r >= J.mi_rows || c >= J.mi_cols) return false;
        auto skip_at = [&](uint32_t rr, uint32_t cc) -> bool {
            return (gread(J.lf_mi, (size_t)rr * J.mi_cols + cc) >> OXAV1_MI_SKIP_BIT) & 1u;
        };
        const bool skip = skip_at(r, c) && (r + 1 >= J.mi_rows || skip_at(r + 1, c)) &&
                          (c + 1 >= J.mi_cols || skip_at(r, c + 1)) &&
                          (r + 1 >= J.mi_rows || c + 1 >= J.mi_cols || skip_at(r + 1, c + 1));
        return !skip;
    };
    {
        const int b = tid & 63, quarter = tid >> 6;  // a wave per pair of directions
        if (quarter < 4 && block_active(b))
            cdef_dir_costs<L, true>((const S3<L> *)t0, ls0, lh0, lw0, 8 * (b >> 3), 8 * (b & 7), coeff_shift, quarter,
                              (S3<int32_t> *)(costs + b * 8));
    }
    __syncthreads();
    if (tid < 64) {
        CdefBlk b{};
        if (block_active(tid)) {
            {
                int32_t ydir, var;
                cdef_pick_dir(costs + tid * 8, &ydir, &var);
                const int32_t pri_y = (int32_t)J.cdef.y_pri_strength[idx] << coeff_shift;
                const int32_t sec_y = (int32_t)J.cdef.y_sec_strength[idx] << coeff_shift;
                const int32_t var_str = (var >> 6) > 0 ? min((int32_t)floor_log2((uint32_t)(var >> 6)), 12) : 0;
                b.active = 1;
                b.dir_y = (int8_t)(pri_y == 0 ? 0 : ydir);
                b.pri_y = (int16_t)(var != 0 ? (pri_y * (4 + var_str) + 8) >> 4 : 0);
                b.sec_y = (int16_t)sec_y;
                const int32_t pri_uv = (int32_t)J.cdef.uv_pri_strength[idx] << coeff_shift;
                b.pri_uv = (int16_t)pri_uv;
                b.sec_uv = (int16_t)((int32_t)J.cdef.uv_sec_strength[idx] << coeff_shift);
                b.dir_uv = (int8_t)(pri_uv == 0 ? 0 : c_cdef_uv_dir[J.sx ? 1 : 0][J.sy ? 1 : 0][ydir]);
            }
        }
        blk[tid] = b;
    }
    // luma filter (the taps of block tid are resolved by thread tid, which
    // wrote blk[tid]: no barrier between; tabs->costs, which the taps
    // overwrite, were read by the same wave before)
    {
        const int32_t damping = (int32_t)J.cdef.cdef_damping + coeff_shift;
        const bool inside = ox0 >= kCdefB && oy0 >= kCdefB && ox0 + tw0 + kCdefB <= pw0 && oy0 + th0 + kCdefB <= ph0;
        const bool packed = kCdefPacked<T> && inside && !(tw0 & 1) && !(ls0 & 1) && !(J.cdf_pitch[0] & 1);
        if (tid < 64) {
            const CdefTaps tp = cdef_taps(blk[tid], false, damping, coeff_shift, ls0);
            if (packed)
                tabs->pk[tid] = cdef_taps_pk(tp);
            else
                taps[tid] = tp;
        }
        __syncthreads();
        G<T> *dst = plane_ptr<T>(J.cdf[0]);
        if constexpr (kCdefPacked<T>)
            if (packed) {
                const bool pri = idx >= 0 && J.cdef.y_pri_strength[idx] != 0;
                const bool sec = idx >= 0 && J.cdef.y_sec_strength[idx] != 0;
                cdef_filter_plane_pk_s<T>(pri, sec, t0, ls0, tw0, th0, ox0, oy0, 3, 3, tabs->pk, 6, dst, J.cdf_pitch[0]);
                goto chroma;
            }
        if (inside)
            cdef_filter_plane<L, T, false>(t0, ls0, tw0, th0, ox0, oy0, pw0, ph0, 3, 3, taps, 6, dst, J.cdf_pitch[0]);
        else
            cdef_filter_plane<L, T, true>(t0, ls0, tw0, th0, ox0, oy0, pw0, ph0, 3, 3, taps, 6, dst, J.cdf_pitch[0]);
    }
chroma:
    if (J.np < 2) return;
    // chroma planes (cdef.rs:296-313), each staged in turn
    const int32_t sx = J.sx, sy = J.sy;
    const int32_t damping_uv = (int32_t)J.cdef.cdef_damping + coeff_shift - 1;
    for (int p = 1; p < 3; p++) {
        const int32_t ph = (int32_t)J.pad_rows[p], pw = (int32_t)J.pad_cols[p];
        const int32_t eh = (int32_t)J.cdef_ext_rows[p], ew = (int32_t)J.cdef_ext_cols[p];
        const int32_t oy = oy0 >> sy, ox = ox0 >> sx;
        const int tw = min(kCdefUnit >> sx, ew - ox), th = min(kCdefUnit >> sy, eh - oy);
        const int lw = min(kCdefUnit >> sx, pw - ox), lh = min(kCdefUnit >> sy, ph - oy);
        __syncthreads();
        if (tw <= 0 || th <= 0) continue;
        cdef_load<L, T, kCdefPacked<T>>(t0, lw, lh, ox, oy, pw, ph, plane_ptr<T>(J.deb[p]), J.deb_pitch[p]);
        const int ls = lw + 2 * kCdefB;
        // the Cb and Cr tiles have the same shape, so Cr reuses Cb's taps
        const bool inside = ox >= kCdefB && oy >= kCdefB && ox + tw + kCdefB <= pw && oy + th + kCdefB <= ph;
        const bool packed = kCdefPacked<T> && inside && !(tw & 1) && !(ls & 1) && !(J.cdf_pitch[1] & 1) &&
                            !(J.cdf_pitch[2] & 1);
        if (p == 1 && tid < 64) {
            const CdefTaps tp = cdef_taps(blk[tid], true, damping_uv, coeff_shift, ls);
            if (packed)
                tabs->pk[tid] = cdef_taps_pk(tp);
            else
                taps[tid] = tp;
        }
        __syncthreads();
        G<T> *dst = plane_ptr<T>(J.cdf[p]);
        if constexpr (kCdefPacked<T>)
            if (packed) {
                const bool pri = idx >= 0 && J.cdef.uv_pri_strength[idx] != 0;
                const bool sec = idx >= 0 && J.cdef.uv_sec_strength[idx] != 0;
                cdef_filter_plane_pk_s<T>(pri, sec, t0, ls, tw, th, ox, oy, 3 - sx, 3 - sy, tabs->pk, 6 - sx, dst,
                                          J.cdf_pitch[p]);
                continue;
            }
        if (inside)
            cdef_filter_plane<L, T, false>(t0, ls, tw, th, ox, oy, pw, ph, 3 - sx, 3 - sy, taps, 6 - sx, dst,
                                           J.cdf_pitch[p]);
        else
            cdef_filter_plane<L, T, true>(t0, ls, tw, th, ox, oy, pw, ph, 3 - sx, 3 - sy, taps, 6 - sx, dst,
                                          J.cdf_pitch[p]);
    }
}

}  // namespace oxav1
