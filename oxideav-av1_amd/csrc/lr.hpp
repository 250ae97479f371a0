// §7.17 loop restoration on gfx950: one workgroup per (plane, 64-luma-row
// stripe, 32-row chunk of the stripe, column tile), Wiener or self-guided
// chosen per tile.
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
// :686-699 weights even rows 0, and block rows start even).  A luma stripe is
// split into two 32-row chunks (the routing still follows the stripe), which
// halves the LDS footprint so twice as many workgroups fit on a CU.
#pragma once

#include "common.hpp"

namespace oxav1 {

constexpr int kLrWaves = 4;                 // waves per workgroup (256 threads)
constexpr int kLrMaxW = 64;                 // max tile width (plane samples)
constexpr int kLrMaxH = 32;                 // max rows of one chunk of a stripe
constexpr int kLrSw = kLrMaxW + 8;          // source window row pitch: window columns -1 .. W+6
                                            // (the window's first column at LDS column 1)
constexpr int kLrSh = kLrMaxH + 6;          // source window height
constexpr int kLrAw = kLrMaxW + 2;          // A/B columns: tile cols -1 .. W
constexpr int kLrAh1 = kLrMaxH + 2;         // pass-1 A/B rows: tile rows -1 .. H
constexpr int kLrAh0 = kLrMaxH / 2 + 1;     // pass-0 A/B rows: the odd rows only
// Output rows per wave when four waves share a chunk, and A/B rows per
// thread band when at least three bands of (tile width + 2) columns fit in a
// workgroup: the row loops are unrolled to these bounds, so their register
// windows need no rotating copies.
constexpr int kLrBand = (kLrMaxH + 3) / 4;
static_assert(kLrBand % 2 == 0, "output row steps keep the row parity");
constexpr int kLrAbBand = (kLrMaxH + 2 + 2) / 3;
static_assert(kLrAbBand % 2 == 0, "A/B row steps keep the row parity");

// n / d for n < 2^24 and 1 <= d < 2^16 as one multiply and shift: with
// m = floor(2^40 / d) + 1 = 2^40 / d + e, 0 < e <= 1, n * m / 2^40 is
// n / d + n * e / 2^40, and the second term is below 2^-16 < 1 / d, so the
// floor is the quotient.
__host__ __device__ inline uint64_t lr_magic(uint32_t d) { return (((uint64_t)1 << 40) / (d ? d : 1u)) + 1; }
__device__ __forceinline__ uint32_t lr_mdiv(uint32_t n, uint64_t m) { return (uint32_t)(((uint64_t)n * m) >> 40); }
__constant__ uint64_t c_lr_recip[130];      // lr_magic(k), k = 1..129


__constant__ int32_t c_sgr_params[16][4];   // filled from kSgrParams at init
__constant__ uint16_t c_sgr_a2[256];        // a2 of z = 0 .. 255 (loop_restoration.rs:648-654)


// A (<= 256, 9 bits) at bit 22 and B (< 2^20) below it, of one self-guided
// position in one word.  Two or three words add field by field without a
// carry between the fields (B sums < 3 * 2^20 < 2^22, A sums <= 768 < 2^10),
// so neighbouring positions are summed before the fields are split.
constexpr uint32_t kLrAbShift = 22, kLrAbMask = (1u << kLrAbShift) - 1;
__device__ __forceinline__ uint32_t lr_pack_ab(uint32_t a, uint32_t b) { return (a << kLrAbShift) | b; }

// Chunks per stripe: a luma stripe (64 rows) has two, a vertically
// subsampled chroma stripe (32 rows) one.
constexpr uint32_t lr_chunks(uint32_t sy) { return (64u >> sy) / kLrMaxH; }

template <typename L>
struct LrSmem {
    alignas(16) L src[kLrSh * kLrSw];  // routed source window, rows ys-3 .. ye+3
    union {
        struct {
            uint32_t ab1[kLrAh1 * kLrAw];  // pass-1 (A, B)
            uint32_t ab0[kLrAh0 * kLrAw];  // pass-0 (A, B), odd rows
        };
        int32_t inter[kLrSh * kLrMaxW];    // Wiener horizontal pass
    };
    uint16_t a2[256];                  // c_sgr_a2 staged for lane-indexed reads
};

struct LrTileGeom {
    int32_t ss, se;      // stripe start / end rows (may be < 0 / > plane end)
    int32_t pex, pey;    // PlaneEndX / PlaneEndY
    int32_t ys, ye;      // output rows (inclusive)
    int32_t xs, xe;      // output cols (inclusive)
    int32_t wy0, wx0;    // window origin (virtual coordinates)
    int32_t ww, wh;      // window size
};

// Self-guided A/B from box sums (loop_restoration.rs:615-665): a = sum of
// squares, b = sum over the (2r+1)^2 window.  Returns lr_pack_ab(A, B).
// K24: samples are u8/u16 (the engine's frames), so the bounds below hold
// and the products run on the 24-bit multiplier; the drop-in path's int32
// samples keep 32-bit multiplies.
template <bool K24, int32_t N>
__device__ __forceinline__ uint32_t lr_ab(uint32_t a, int32_t b, uint32_t s, uint32_t bd, const uint16_t *a2tab) {
    constexpr uint32_t kOon = ((1u << 12) + N / 2) / N;  // oneOverN
    const uint32_t sha = 2 * (bd - 8), shb = bd - 8;
    // Round2 by sha / shb (a shift by 0 adds 0)
    const uint32_t ar = (a + ((1u << sha) >> 1)) >> sha;
    const int32_t dr = (b + ((1 << shb) >> 1)) >> shb;
    // p = ar * n - dr^2 clipped at 0 is at most n^2 / 4 squared sample
    // ranges (two-valued windows; 1.311e6 for r = 1, 10.23e6 for r = 2 at
    // every depth, rounding included), and s <= 3236 for r = 1, s <= 140 for
    // r = 2: p * s < 2^32 (4.25e9 at most), as the reference's saturating
    // 64-bit product never saturates.  (256 - a2) * oneOverN <= 255 * 455
    // and b < 2^17, so b2 < 2^32.  K24: ar < 2^21, 0 <= dr <= 25 * 256
    // and p < 2^24, so every product runs on the 24-bit multiplier.
    uint32_t a2, b2;
    if constexpr (K24) {
        int32_t p = (int32_t)mul_u24k<N>(ar) - (int32_t)mul_u24((uint32_t)dr, (uint32_t)dr);
        p = p < 0 ? 0 : p;
        const uint32_t z = (mul_u24_s(s, (uint32_t)p) + (1u << 19)) >> 20;
        // the table holds 1 at z = 0 and 256 at z = 255, so every z >= 255
        // reads its last entry
        a2 = a2tab[z < 255u ? z : 255u];
        b2 = mul_u24(mul_u24k<kOon>(256u - a2), (uint32_t)b);
    } else {
        int32_t p = (int32_t)ar * N - dr * dr;
        p = p < 0 ? 0 : p;
        const uint32_t z = ((uint32_t)p * s + (1u << 19)) >> 20;
        a2 = a2tab[z < 255u ? z : 255u];
        b2 = (256u - a2) * kOon * (uint32_t)b;
    }
    return lr_pack_ab(a2, (b2 + 2048u) >> 12);
}

// s = (2^20 + n^2 e / 2) / (n^2 e) of one self-guided pass
// (loop_restoration.rs:612-616) per (set, pass): it depends on Sgr_Params
// only, so the host computes it once at device set-up instead of every
// thread per work item (a 64-bit division each); n and oneOverN follow from
// the radius, which the kernels know at compile time.
__constant__ uint32_t c_sgr_s[16][2];

__host__ __device__ inline uint32_t lr_sgr_s_of(int32_t r, int32_t e) {
    if (r == 0) return 0;
    const int64_t n = (2 * r + 1) * (2 * r + 1);
    const int64_t n2e = n * n * e;
    return (uint32_t)((((int64_t)1 << 20) + n2e / 2) / (n2e > 1 ? n2e : 1));
}

// unit(): the tile's restoration unit (called once the unit is resolved);
// put(): a copied or Wiener sample; sgr(): a self-guided sample with its
// filter outputs (u = the CDEF sample, f0/f1 = the two passes' flt values,
// u << 4 where a pass is off).
// Four consecutive samples of T as one vector access.
typedef uint32_t lr_u32x2 __attribute__((ext_vector_type(2)));
typedef int32_t lr_i32x4 __attribute__((ext_vector_type(4)));
template <typename T>
struct LrVec;
template <>
struct LrVec<uint8_t> {
    using type = uint32_t;
};
template <>
struct LrVec<uint16_t> {
    using type = lr_u32x2;
};
template <>
struct LrVec<int32_t> {
    using type = lr_i32x4;
};
// the four samples of v into LDS at d (8-byte aligned for u16, 16 for int32)
template <typename T, typename L>
__device__ __forceinline__ void lr_put4(L *d, typename LrVec<T>::type v) {
    if constexpr (std::is_same<T, uint8_t>::value) {
        *(lr_u32x2 *)d = lr_u32x2{(v & 0xffu) | ((v << 8) & 0xff0000u), ((v >> 16) & 0xffu) | ((v >> 8) & 0xff0000u)};
    } else {
        *(typename LrVec<T>::type *)d = v;
    }
}

// Where a tile's restored samples go.  LrStore writes them to the output
// plane; the encoder's election passes sinks that reduce them instead.
// kCopy4: the sink takes a NONE tile's copy four samples at a time (copy4).
template <typename T>
struct LrStore {
    static constexpr bool kCopy4 = true;
    G<T> *__restrict__ out;
    uint32_t opitch;
    __device__ __forceinline__ bool copy4_ok() const {
        return (opitch & 3u) == 0 && ((uintptr_t)out & (4 * sizeof(T) - 1)) == 0;
    }
    // offsets in 32 bits: rows x pitch of an output plane stays below 2^31
    // (checked when the job is built)
    __device__ __forceinline__ void copy4(int32_t y, int32_t x, typename LrVec<T>::type v) {
        *(G<typename LrVec<T>::type> *)(out + ((uint32_t)y * opitch + (uint32_t)x)) = v;
    }
    __device__ __forceinline__ void unit(uint32_t, uint32_t, const oxav1_lr_unit &) {}
    __device__ __forceinline__ void put(int32_t y, int32_t x, int32_t v) {
        out[(uint32_t)y * opitch + (uint32_t)x] = (T)v;
    }
    __device__ __forceinline__ void sgr(int32_t y, int32_t x, int32_t v, int32_t, int32_t, int32_t) {
        out[(uint32_t)y * opitch + (uint32_t)x] = (T)v;
    }
};

// (stripe, chunk, column tile) of item t of plane p's loop-restoration items,
// numbered stripe-major, then chunk, then tile.
template <class JT>
__device__ __forceinline__ void lr_decode(JT &J, uint32_t p, uint32_t t, uint32_t &stripe, uint32_t &chunk,
                                          uint32_t &tile) {
    const uint32_t q = lr_mdiv(t, J.lr_magic_tx[p]);  // stripe * chunks + chunk
    tile = t - q * J.lr_tiles_x[p];
    const bool two = lr_chunks(p ? J.sy : 0) == 2;
    stripe = two ? q >> 1 : q;
    chunk = two ? q & 1u : 0u;
}

// One tile: (plane, stripe, column tile) -> sink (cropped frame coordinates
// equal plane coordinates).  Emits rows [ys, ye] x cols [xs, xe].  JT is the
// job record's type: the engine passes it in the constant address space, so
// its fields (uniform in the workgroup) become scalar loads.
template <typename T, class Sink, class JT = const FrameJob>
__device__ __forceinline__ void lr_tile(JT &J, uint32_t plane, uint32_t stripe, uint32_t chunk, uint32_t tile,
                        LrSmem<typename LdsOf<T>::type> &S, Sink &sink) {
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
    // writes are dropped outside the output buffer (loop_restoration.rs:814-818)
    g.ye = min(g.ye, (int32_t)J.out_rows[plane] - 1);
    g.xe = min(g.xe, (int32_t)J.out_cols[plane] - 1);
    if (g.ys > g.ye || g.xs > g.xe) return;
    const int tid = threadIdx.x, nt = blockDim.x;
    // the wave index is uniform in a wave: a scalar, so row bands and their
    // base addresses are scalar too
    const int lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    constexpr int nwv = kLrWaves;
    const int32_t W = g.xe - g.xs + 1;
    const G<T> *cdf = plane_ptr<T>(J.up_cdef[plane]);
    const uint32_t dpitch = J.up_cdef_pitch[plane];

    // unit of this tile (loop_restoration.rs:395-405).  lr_type is read as
    // the aligned word holding lr_type[0..2]: a byte load indexed by the
    // plane lets the compiler form the address J + plane and reuse it as the
    // base of scalar loads of the dword arrays below (J + plane + 3 * plane +
    // offset), and a scalar load ignores the low two bits of its base, so
    // those loads would read another plane's fields.
    const uint32_t types = J.lr_type4;
    uint32_t type = J.uses_lr ? (types >> (8 * plane)) & 0xffu : OXAV1_RESTORE_NONE;
    oxav1_lr_unit u{};
    if (type != OXAV1_RESTORE_NONE) {
        const uint64_t mus = J.lr_magic_us[plane];
        const uint32_t ur = min(J.lr_unit_rows[plane] - 1, lr_mdiv((uint32_t)(g.ys + (8 >> sy)), mus));
        const uint32_t uc = min(J.lr_unit_cols[plane] - 1, lr_mdiv((uint32_t)g.xs, mus));
        u = J.lr_units[plane] ? J.lr_units[plane][(size_t)ur * J.lr_unit_cols[plane] + uc] : oxav1_lr_unit{};
        type = u.type;
        sink.unit(ur, uc, u);
    }
    // this chunk's rows of the stripe
    g.ys += (int32_t)chunk * kLrMaxH;
    g.ye = min(g.ye, g.ys + kLrMaxH - 1);
    if (g.ys > g.ye) return;
    const int32_t H = g.ye - g.ys + 1;
    // Output rows are split into one band per wave; a lane owns a column.
    // (an even count: the first rows ys and b0 are then even, and so is
    // every row step, so a row's parity is its unrolled position's)
    const int RB = (((H + nwv - 1) >> 2) + 1) & ~1;
    const int b0 = wv * RB, b1 = min(H, b0 + RB);
    using V = typename LrVec<T>::type;
    if (type != OXAV1_RESTORE_WIENER && type != OXAV1_RESTORE_SGRPROJ) {
        if (!J.lr_copy_none) return;
        if constexpr (Sink::kCopy4) {
          if (sink.copy4_ok() && (dpitch & 3u) == 0 && ((uintptr_t)cdf & (4 * sizeof(T) - 1)) == 0) {
            // four samples per lane (the tile starts at a multiple of 32):
            // thread t copies group t % 16 of rows t / 16 and t / 16 + 16 (a
            // tile is at most 16 groups wide and 32 rows high), both loads
            // issued before either store
            static_assert(kLrMaxW == 64 && kLrMaxH == 32 && kLrWaves == 4, "copy mapping");
            const int gi = tid & 15, r = tid >> 4;
            const int32_t x = g.xs + 4 * gi;
            if (x > g.xe) return;
            const G<T> *q0 = cdf + (size_t)(g.ys + r) * dpitch + x;
            const G<T> *q1 = q0 + (size_t)16 * dpitch;
            const bool has0 = r < H, has1 = r + 16 < H;
            if (x + 3 <= g.xe) {
                V v0{}, v1{};
                if (has0) v0 = *(const G<V> *)q0;
                if (has1) v1 = *(const G<V> *)q1;
                if (has0) sink.copy4(g.ys + r, x, v0);
                if (has1) sink.copy4(g.ys + r + 16, x, v1);
            } else {
                for (int k = 0; k < 4 && x + k <= g.xe; k++) {
                    if (has0) sink.put(g.ys + r, x + k, (int32_t)q0[k]);
                    if (has1) sink.put(g.ys + r + 16, x + k, (int32_t)q1[k]);
                }
            }
            return;
          }
        }
        // every load of the band is issued before the first store, so the
        // copy waits for memory once, not once per row
        for (int x = lane; x < W; x += 64)
            for (int rb = b0; rb < b1; rb += kLrBand) {
                T v[kLrBand];
#pragma unroll
                for (int i = 0; i < kLrBand; i++)
                    v[i] = cdf[(size_t)(g.ys + min(rb + i, b1 - 1)) * dpitch + (g.xs + x)];
#pragma unroll
                for (int i = 0; i < kLrBand; i++)
                    if (rb + i < b1) sink.put(g.ys + rb + i, g.xs + x, (int32_t)v[i]);
            }
        return;
    }
    // Routed source window (get_source_sample, loop_restoration.rs:862-902):
    // rows above/below the stripe come from the pre-CDEF frame (at most 2
    // rows out), the rest from the CDEF frame; then both clamps.  The window
    // (rows ys-3 .. ye+3, columns xs-3 .. xe+3) is loaded in groups of four
    // columns starting at xs-4 (a multiple of four): window column c sits at
    // LDS column c + 1, so every group is one aligned LDS store.  A lane
    // routes its group's row and loads the four samples with one vector
    // load where they lie inside the row's buffer, clamped per sample at the
    // plane's edges.  All of a lane's groups are loaded before the first
    // store (one memory round trip); rows without a buffer store 0.
    g.wy0 = g.ys - 3;
    g.wx0 = g.xs - 3;
    g.wh = H + 6;
    g.ww = W + 6;
    if (type == OXAV1_RESTORE_SGRPROJ)
        for (int i = tid; i < 256; i += nt) S.a2[i] = c_sgr_a2[i];
    {
        const int32_t crows = (int32_t)J.lrc_rows[plane], ccols = (int32_t)J.lrc_cols[plane];
        const int32_t rows = (int32_t)J.lr_rows[plane], cols = (int32_t)J.lr_cols[plane];
        const G<T> *deb = plane_ptr<T>(J.up_deb[plane]);
        const uint32_t bpitch = J.up_deb_pitch[plane];
        const bool al_c = (dpitch & 3u) == 0 && ((uintptr_t)cdf & (4 * sizeof(T) - 1)) == 0;
        const bool al_d = (bpitch & 3u) == 0 && ((uintptr_t)deb & (4 * sizeof(T) - 1)) == 0;
        const int32_t xmax = max(g.pex, 0);
        const int ngx = (g.ww + 1 + 3) >> 2;  // groups per row: window columns -1 .. ww-1
        const uint64_t mg = c_lr_recip[ngx];
        const int ng = g.wh * ngx;
        constexpr int kG = (kLrSh * (kLrSw / 4) + 255) / 256;  // groups per lane
        V v[kG];
        int off[kG];
#pragma unroll
        for (int j = 0; j < kG; j++) {
            const int i = tid + j * 256;
            off[j] = -1;
            if (i < ng) {
                const int r = (int)lr_mdiv((uint32_t)i, mg), gi = i - r * ngx;
                int32_t y = clip3(0, max(g.pey, 0), g.wy0 + r);
                const bool alt = y < g.ss || y > g.se;
                if (alt) y = y < g.ss ? max(y, g.ss - 2) : min(y, g.se + 2);
                const int32_t nr = alt ? crows : rows, nc = alt ? ccols : cols;
                const G<T> *b = alt ? deb : cdf;
                const uint32_t pitch = alt ? bpitch : dpitch;
                off[j] = r * kLrSw + 4 * gi;
                if (nr <= 0 || nc <= 0) {
                    v[j] = V{};
                } else {
                    const G<T> *row = b + (size_t)clip3(0, nr - 1, y) * pitch;
                    const int32_t x0 = g.xs - 4 + 4 * gi, cm = min(xmax, nc - 1);
                    if (x0 >= 0 && x0 + 3 <= cm && (alt ? al_d : al_c)) {
                        v[j] = *(const G<V> *)(row + x0);
                    } else {
                        T t[4];
#pragma unroll
                        for (int k = 0; k < 4; k++) t[k] = row[min(clip3(0, xmax, x0 + k), nc - 1)];
                        if constexpr (std::is_same<T, uint8_t>::value)
                            v[j] = (uint32_t)t[0] | (uint32_t)t[1] << 8 | (uint32_t)t[2] << 16 | (uint32_t)t[3] << 24;
                        else if constexpr (std::is_same<T, uint16_t>::value)
                            v[j] = V{(uint32_t)t[0] | (uint32_t)t[1] << 16, (uint32_t)t[2] | (uint32_t)t[3] << 16};
                        else
                            v[j] = V{t[0], t[1], t[2], t[3]};
                    }
                }
            }
        }
#pragma unroll
        for (int j = 0; j < kG; j++)
            if (off[j] >= 0) lr_put4<T>(S.src + off[j], v[j]);
    }
    __syncthreads();
    const int32_t bd = J.bd;
    const int32_t maxv = (1 << bd) - 1;
    if (type == OXAV1_RESTORE_WIENER) {
        // wiener_filter (loop_restoration.rs:742-821); both 7-tap filters are
        // symmetric, so taps t and 6-t share one multiply; the taps are the
        // unit's (wave-uniform), multiplied from scalar registers
        int32_t vf[4], hf[4];
        vf[3] = hf[3] = 128;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            vf[i] = u.wiener[0][i];
            hf[i] = u.wiener[1][i];
            vf[3] -= 2 * u.wiener[0][i];
            hf[3] -= 2 * u.wiener[1][i];
        }
        const uint32_t r0 = bd == 12 ? 5 : 3, r1 = bd == 12 ? 9 : 11;
        const int32_t offset = 1 << (bd + 7 - (int32_t)r0 - 1);
        const int32_t limit = (1 << (bd + 1 + 7 - (int32_t)r0)) - 1;
        // horizontal pass: window rows 0 .. H+5 (tile rows -3 .. H+2) into
        // LDS, one row per wave at a time, lanes on columns
        for (int rb = 0; rb < H + 6; rb += nwv) {
            const int wr = rb + wv;
            if (wr < H + 6 && lane < W) {
                const L *q = S.src + 1 + wr * kLrSw + lane;
                // u16 samples: |taps| < 2^9 and sample sums < 2^17, 24-bit multiplies
                constexpr bool k24h = std::is_same<L, uint16_t>::value;
                auto hm = [](int32_t x, int32_t y, int32_t z) -> int32_t { return k24h ? mad_i24_s(x, y, z) : x * y + z; };
                const int32_t acc = hm(hf[0], (int32_t)q[0] + (int32_t)q[6],
                                       hm(hf[1], (int32_t)q[1] + (int32_t)q[5],
                                          hm(hf[2], (int32_t)q[2] + (int32_t)q[4],
                                             hm(hf[3], (int32_t)q[3], 1 << (r0 - 1)))));
                S.inter[wr * kLrMaxW + lane] = med3_i32_s(acc >> r0, -offset, limit - offset);
            }
        }
        __syncthreads();
        if (lane >= W || b0 >= b1) return;
        // vertical pass down the wave's band, kLrBand rows at a time from
        // registers.  The horizontal results are clipped to
        // [-offset, limit - offset] (|h| < 2^15) and |vf| < 2^18 for any
        // int16 taps, so the 24-bit multiplies give the 32-bit products.
        const int32_t *iv = S.inter + lane;
        for (int rb = b0; rb < b1; rb += kLrBand) {
            int32_t h[kLrBand + 6];
#pragma unroll
            for (int i = 0; i < kLrBand + 6; i++) h[i] = rb + i < H + 6 ? iv[(rb + i) * kLrMaxW] : 0;
#pragma unroll
            for (int i = 0; i < kLrBand; i++)
                if (rb + i < b1) {
                    const int32_t acc =
                        mad_i24_s(vf[0], h[i] + h[i + 6],
                                  mad_i24_s(vf[1], h[i + 1] + h[i + 5],
                                            mad_i24_s(vf[2], h[i + 2] + h[i + 4], mad_i24_s(vf[3], h[i + 3], 1 << (r1 - 1)))));
                    sink.put(g.ys + rb + i, g.xs + lane, clamp0_s(acc >> r1, maxv));
                }
        }
        return;
    }
    // self_guided_filter (loop_restoration.rs:485-579) + box_filter (594-729)
    // The pass radii (r0 in {0, 2}, r1 in {0, 1}, never both 0) are
    // compile-time in the body below: one instance per radius pair, so the
    // unrolled row loops carry no per-row radius tests, and n and oneOverN
    // are constants.
    const uint32_t set = u.sgr_set < 15 ? u.sgr_set : 15;
    const int32_t w0 = u.sgr_xqd[0], w1 = u.sgr_xqd[1];
    const int32_t w2 = 128 - w0 - w1;
    auto body = [&](auto r0c, auto r1c) {
    constexpr int32_t r0 = decltype(r0c)::value, r1 = decltype(r1c)::value;
    constexpr int32_t n0 = (2 * r0 + 1) * (2 * r0 + 1), n1 = (2 * r1 + 1) * (2 * r1 + 1);
    // A/B over tile rows -1 .. H and cols -1 .. W: thread -> (column, row
    // band); each walks down its band keeping 3- and 5-wide horizontal sums
    // of the last five window rows, so every box sum costs five LDS reads.
    // Pass 0 (radius r0 = 2 when used) is needed on odd rows only: its
    // even-row output weights are 0 and block rows start even
    // (loop_restoration.rs:686-699).
    {
        const int ncol = W + 2;
        const uint64_t mc = c_lr_recip[ncol];
        const int nb = (int)lr_mdiv((uint32_t)nt, mc);
        const int band = (int)lr_mdiv((uint32_t)tid, mc), c = tid - band * ncol;
        // rows per band, rounded up to even: every band then starts on an odd
        // tile row, so a row's parity follows from its unrolled position
        const int KR = ((int)lr_mdiv((uint32_t)(H + 2 + nb - 1), c_lr_recip[nb]) + 1) & ~1;
        const int k0 = -1 + band * KR, k1 = min(H, k0 + KR - 1);  // inclusive tile rows
        if (band < nb && k0 <= k1) {
            const uint32_t s0 = c_sgr_s[set][0], s1 = c_sgr_s[set][1];
            // window row of tile row k is k + 3; column of tile col c - 1 is c + 2
            int32_t h3[5], q3[5], h5[5], q5[5];
            auto hsum = [&](int wr, int slot) {
                const L *q = S.src + 1 + wr * kLrSw + c;  // cols c .. c+4 = tile cols c-3 .. c+1
                const int32_t v0 = q[0], v1 = q[1], v2 = q[2], v3 = q[3], v4 = q[4];
                h3[slot] = v1 + v2 + v3;
                q3[slot] = v1 * v1 + v2 * v2 + v3 * v3;
                h5[slot] = h3[slot] + v0 + v4;
                q5[slot] = q3[slot] + v0 * v0 + v4 * v4;
            };
            // The band is walked kLrAbBand rows at a time (one step when
            // three or more bands fit, W + 2 <= 85), each step unrolled so
            // the five-row windows are renamed, not copied: rows kb-2 ..
            // kb+1 preloaded into slots 0..3, then one new row per output row.
            const int32_t ys = g.ys;
            for (int kb = k0; kb <= k1; kb += kLrAbBand) {
#pragma unroll
                for (int j = 0; j < 4; j++) hsum(kb - 2 + j + 3, j);
#pragma unroll
                for (int i = 0; i < kLrAbBand; i++) {
                    const int k = kb + i;
                    if (k > k1) break;
                    hsum(k + 2 + 3, 4);
                    if constexpr (r1 != 0) {
                        uint32_t a;
                        int32_t b;
                        if constexpr (r1 == 1) {
                            a = (uint32_t)(q3[1] + q3[2] + q3[3]);
                            b = h3[1] + h3[2] + h3[3];
                        } else {
                            a = (uint32_t)q5[0] + (uint32_t)q5[1] + (uint32_t)q5[2] + (uint32_t)q5[3] + (uint32_t)q5[4];
                            b = h5[0] + h5[1] + h5[2] + h5[3] + h5[4];
                        }
                        S.ab1[(k + 1) * kLrAw + c] = lr_ab<std::is_same<L, uint16_t>::value, n1>(a, b, s1, (uint32_t)bd, S.a2);
                    }
                    // (ys + k) & 1 with k = kb + i and kb odd (k0 is -1 plus
                    // an even number of rows, and steps are even)
                    if (r0 != 0 && ((ys + 1 + i) & 1)) {
                        const uint32_t a =
                            (uint32_t)q5[0] + (uint32_t)q5[1] + (uint32_t)q5[2] + (uint32_t)q5[3] + (uint32_t)q5[4];
                        const int32_t b = h5[0] + h5[1] + h5[2] + h5[3] + h5[4];
                        S.ab0[((k + 1) >> 1) * kLrAw + c] = lr_ab<std::is_same<L, uint16_t>::value, n0>(a, b, s0, (uint32_t)bd, S.a2);
                    }
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        h3[j] = h3[j + 1];
                        q3[j] = q3[j + 1];
                        h5[j] = h5[j + 1];
                        q5[j] = q5[j + 1];
                    }
                }
            }
        }
    }
    __syncthreads();
    if (lane >= W || b0 >= b1) return;
    // Weighted 3x3 of A/B (loop_restoration.rs:686-729), walking down the
    // band: pass 1 weights 3 on corners and 4 elsewhere, i.e. for rows
    // k-1, k+1: 3*S + M, row k: 4*S (S = three-column sum, M = centre);
    // pass 0 (odd rows R hold P(R) = 5*(left + right) + 6*centre): odd rows
    // use P(k) >> 4, even rows (P(k-1) + P(k+1)) >> 5.
    const int32_t ys = g.ys;
    auto row3 = [&](const uint32_t *ab, int idx, int32_t &sa, int32_t &ma, int32_t &sb, int32_t &mb) {
        const uint32_t l = ab[idx + lane], m = ab[idx + lane + 1], r = ab[idx + lane + 2];
        const uint32_t s3 = l + m + r;  // field by field (lr_pack_ab)
        ma = (int32_t)(m >> kLrAbShift);
        mb = (int32_t)(m & kLrAbMask);
        sa = (int32_t)(s3 >> kLrAbShift);
        sb = (int32_t)(s3 & kLrAbMask);
    };
    auto prow = [&](int k, int32_t &pa, int32_t &pb) {  // P at odd tile row k
        const int idx = ((k + 1) >> 1) * kLrAw;
        const uint32_t l = S.ab0[idx + lane], m = S.ab0[idx + lane + 1], r = S.ab0[idx + lane + 2];
        const uint32_t lr2 = l + r;  // field by field (lr_pack_ab)
        pa = 5 * (int32_t)(lr2 >> kLrAbShift) + 6 * (int32_t)(m >> kLrAbShift);
        pb = 5 * (int32_t)(lr2 & kLrAbMask) + 6 * (int32_t)(m & kLrAbMask);
    };
    int32_t sa0 = 0, ma0 = 0, sb0 = 0, mb0 = 0, sa1 = 0, ma1 = 0, sb1 = 0, mb1 = 0;
    if constexpr (r1 != 0) {
        row3(S.ab1, (b0 - 1 + 1) * kLrAw, sa0, ma0, sb0, mb0);
        row3(S.ab1, (b0 + 1) * kLrAw, sa1, ma1, sb1, mb1);
    }
    // pass-0 P of the odd row at or above the current row (tile row -1 is odd)
    int32_t pa_prev = 0, pb_prev = 0;
    if constexpr (r0 != 0) {
        const int kp = ((ys + b0) & 1) ? b0 : b0 - 1;
        prow(kp, pa_prev, pb_prev);
    }
    // With 16-bit samples every product below has both operands under 2^23
    // (A sums <= 8192, samples < 2^16, filter outputs < 2^22, weights from
    // int16 xqd), so it runs on the 24-bit multiplier.
    constexpr bool k24 = std::is_same<L, uint16_t>::value;
    auto mad = [](int32_t x, int32_t y, int32_t z) -> int32_t { return k24 ? mad_i24(x, y, z) : x * y + z; };
    // w0..w2 are the unit's (wave-uniform) weights
    auto madw = [](int32_t w, int32_t y, int32_t z) -> int32_t { return k24 ? mad_i24_s(w, y, z) : w * y + z; };
    for (int kb = b0; kb < b1; kb += kLrBand) {
#pragma unroll
        for (int i = 0; i < kLrBand; i++) {
            const int k = kb + i;
            if (k >= b1) break;
            const int32_t uval = (int32_t)S.src[1 + (k + 3) * kLrSw + lane + 3];
            const int32_t u4 = uval << 4;
            int32_t f0 = u4, f1 = u4;
            if constexpr (r0 != 0) {
                int32_t as, bs;
                uint32_t sh;
                // (ys + k) & 1: ys, kb and the steps are even
                if (i & 1) {  // P(k) is already held: set at b0 or by row k-1
                    as = pa_prev;
                    bs = pb_prev;
                    sh = 4;
                } else {
                    int32_t pa_next, pb_next;
                    prow(k + 1, pa_next, pb_next);
                    as = pa_prev + pa_next;
                    bs = pb_prev + pb_next;
                    pa_prev = pa_next;
                    pb_prev = pb_next;
                    sh = 5;
                }
                const int32_t v = mad(as, uval, bs);
                const uint32_t ts = 8 + sh - 4;
                f0 = (v + (1 << (ts - 1))) >> ts;
            }
            if constexpr (r1 != 0) {
                int32_t sa2, ma2, sb2, mb2;
                row3(S.ab1, (k + 2) * kLrAw, sa2, ma2, sb2, mb2);
                const int32_t as = 3 * (sa0 + sa2) + ma0 + ma2 + 4 * sa1;
                const int32_t bs = 3 * (sb0 + sb2) + mb0 + mb2 + 4 * sb1;
                const int32_t v = mad(as, uval, bs);
                f1 = (v + (1 << 8)) >> 9;
                sa0 = sa1, ma0 = ma1, sb0 = sb1, mb0 = mb1;
                sa1 = sa2, ma1 = ma2, sb1 = sb2, mb1 = mb2;
            }
            const int32_t v = madw(w2, r1 != 0 ? f1 : u4, madw(w0, r0 != 0 ? f0 : u4, madw(w1, u4, 1 << 10)));
            const int32_t sres = v >> 11;
            sink.sgr(ys + k, g.xs + lane, clamp0_s(sres, maxv), uval, f0, f1);
        }
    }
    };
    using I0 = std::integral_constant<int32_t, 0>;
    using I1 = std::integral_constant<int32_t, 1>;
    using I2 = std::integral_constant<int32_t, 2>;
    const int32_t r0 = c_sgr_params[set][0], r1 = c_sgr_params[set][2];
    if (r0 == 0)
        body(I0{}, I1{});
    else if (r1 == 0)
        body(I2{}, I0{});
    else
        body(I2{}, I1{});
}

}  // namespace oxav1
