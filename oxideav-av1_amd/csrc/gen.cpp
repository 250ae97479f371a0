// Seeded synthetic workload generator (see include/oxav1_gen.h).
//
// Produces legal pre-filter frames, per-mi grids and filter parameters in
// the ranges the reference's parsers accept:
//   LoopFilterParams  uncompressed_header_tail.rs:261-291 (levels 0..63,
//                     sharpness 0..7, default ref deltas :239)
//   CdefParams        uncompressed_header_tail.rs:876-900 (damping 3..6,
//                     sec strengths {0,1,2,4})
//   LR units          cdf.rs:20055-20107 (Wiener taps in [MIN,MAX], chroma
//                     tap 0 forced to 0; SGR xqd with the r==0 overrides)
//   FilmGrainParams   uncompressed_header_tail.rs:1763-1834
// Frame geometry: frame_header.rs:1742-1746.  Chroma loop-filter transform
// sizes follow the walker bridge (cdf.rs:20559-20594).
#include "../../include/oxav1_gen.h"

#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "av1_tables.h"

namespace {

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    // uniform in [lo, hi]
    int32_t range(int32_t lo, int32_t hi) {
        uint64_t span = (uint64_t)((int64_t)hi - (int64_t)lo + 1);
        return (int32_t)((int64_t)lo + (int64_t)(next() % span));
    }
    bool chance(uint32_t num, uint32_t den) { return (next() % den) < num; }
};

// BLOCK ordinal for (w4, h4) in 4-sample units; -1 if none.
int block_of(uint32_t w4, uint32_t h4) {
    for (int b = 0; b < 22; b++)
        if (oxav1::kNum4x4W[b] == w4 && oxav1::kNum4x4H[b] == h4) return b;
    return -1;
}
// TX ordinal for (w, h) in samples; -1 if none.
int tx_of(uint32_t w, uint32_t h) {
    for (int t = 0; t < 19; t++)
        if (oxav1::kTxWidth[t] == w && oxav1::kTxHeight[t] == h) return t;
    return -1;
}

// Split_Tx_Size (cdf.rs:12922-12942) by its closed form: a square
// transform halves both sides (TX_4X4 stays), a rectangular one halves its
// longer side.
int split_tx(int tx) {
    uint32_t w = oxav1::kTxWidth[tx], h = oxav1::kTxHeight[tx];
    if (w == h) return w == 4 ? tx : tx_of(w / 2, h / 2);
    return w > h ? tx_of(w / 2, h) : tx_of(w, h / 2);
}

int residual_size(int subsize, int plane, int sx, int sy) {
    if (subsize < 0 || subsize >= 22) return -1;
    int subx = plane > 0 ? (sx != 0) : 0, suby = plane > 0 ? (sy != 0) : 0;
    int raw = oxav1::kSubsampledSize[subsize * 4 + subx * 2 + suby];
    return raw == OXAV1_BLOCK_INVALID ? -1 : raw;
}

// get_tx_size (cdf.rs:12976-13012); -1 = None.
int chroma_tx(int plane, int tx, int mi_size, int sx, int sy) {
    if (plane == 0) return tx;
    int ps = residual_size(mi_size, plane, sx, sy);
    if (ps < 0) return -1;
    int uv = oxav1::kMaxTxSizeRect[ps];
    int w = oxav1::kTxWidth[uv], h = oxav1::kTxHeight[uv];
    if (w == 64 || h == 64) {
        if (w == 16) return 9;
        if (h == 16) return 10;
        return 3;
    }
    return uv;
}

struct Builder {
    const oxav1_gen_config &cfg;
    Rng rng;
    oxav1_gen_frame *f;
    uint32_t mi_rows, mi_cols;
    std::vector<uint8_t> luma_tx;  // per mi

    Builder(const oxav1_gen_config &c, uint64_t seed, oxav1_gen_frame *fr)
        : cfg(c), rng(seed), f(fr), mi_rows(fr->desc.mi_rows), mi_cols(fr->desc.mi_cols),
          luma_tx((size_t)fr->desc.mi_rows * fr->desc.mi_cols, 0) {}

    // Stamp transform `tx` over the (h4 x w4) mi footprint at (r, c),
    // clipped to the frame.
    void stamp_tx(uint32_t r, uint32_t c, uint32_t w4, uint32_t h4, int tx) {
        for (uint32_t y = r; y < r + h4 && y < mi_rows; y++)
            for (uint32_t x = c; x < c + w4 && x < mi_cols; x++)
                luma_tx[(size_t)y * mi_cols + x] = (uint8_t)tx;
    }

    // read_var_tx_size (cdf.rs:28720-28790): a split tree of at most
    // MAX_VARTX_DEPTH = 2 levels below `tx`, each node split with p = 1/2.
    void var_tx(uint32_t r, uint32_t c, int tx, int depth) {
        uint32_t w4 = oxav1::kTxWidth[tx] / 4, h4 = oxav1::kTxHeight[tx] / 4;
        if (tx != 0 && depth < 2 && rng.chance(1, 2)) {
            int sub = split_tx(tx);
            uint32_t sw = oxav1::kTxWidth[sub] / 4, sh = oxav1::kTxHeight[sub] / 4;
            for (uint32_t i = 0; i < h4; i += sh)
                for (uint32_t j = 0; j < w4; j += sw) var_tx(r + i, c + j, sub, depth + 1);
        } else {
            stamp_tx(r, c, w4, h4, tx);
        }
    }

    void leaf(uint32_t r, uint32_t c, uint32_t w4, uint32_t h4) {
        int bs = block_of(w4, h4);
        int tx = 0;
        if (!cfg.decoder_geometry) {
            // transform: largest rect (<= 32 per side) or a smaller square
            uint32_t bw = w4 * 4, bh = h4 * 4;
            uint32_t mw = bw < 32 ? bw : 32, mh = bh < 32 ? bh : 32;
            if (rng.chance(1, 2)) {
                tx = tx_of(mw, mh);
            } else {
                uint32_t m = mw < mh ? mw : mh;
                uint32_t steps = 0;
                for (uint32_t t = m; t > 4; t >>= 1) steps++;
                uint32_t side = m >> rng.range(0, (int32_t)steps);
                tx = tx_of(side, side);
            }
            if (tx < 0) tx = 0;
        }
        bool skip = rng.chance(3, 10);
        bool inter = rng.chance(1, 2);
        uint8_t ref = inter ? (uint8_t)rng.range(1, 7) : 0;
        uint8_t mode = inter ? (uint8_t)rng.range(14, 25) : (uint8_t)rng.range(0, 12);
        uint8_t seg = cfg.exercise_segments ? (uint8_t)rng.range(0, 7) : 0;
        int8_t dl[4] = {0, 0, 0, 0};
        if (cfg.exercise_delta_lf)
            for (int i = 0; i < 4; i++) dl[i] = (int8_t)rng.range(-20, 20);
        for (uint32_t y = r; y < r + h4 && y < mi_rows; y++)
            for (uint32_t x = c; x < c + w4 && x < mi_cols; x++) {
                size_t k = (size_t)y * mi_cols + x;
                f->mi_size[k] = (uint8_t)bs;
                f->skip[k] = skip;
                f->is_intra[k] = !inter;
                f->ref_frame[k] = ref;
                f->mode[k] = mode;
                f->segment_id[k] = seg;
                for (int i = 0; i < 4; i++) f->delta_lf[k * 4 + i] = dl[i];
                luma_tx[k] = (uint8_t)tx;
            }
        if (cfg.decoder_geometry) {
            // Transform sizes as the decoder's syntax yields them: starting
            // from Max_Tx_Size_Rect (cdf.rs:12870, the 64-sample family
            // included) in tiles of at most 64x64; intra blocks take a
            // uniform tx_depth of 0..2 splits (cdf.rs:28440), skipped inter
            // blocks the maximum, other inter blocks a var-tx split tree.
            int max_tx = oxav1::kMaxTxSizeRect[bs];
            uint32_t tw4 = oxav1::kTxWidth[max_tx] / 4, th4 = oxav1::kTxHeight[max_tx] / 4;
            int uni = max_tx;
            if (!inter) {
                int depth = rng.range(0, 2);
                for (int d = 0; d < depth; d++) uni = split_tx(uni);
            }
            for (uint32_t i = 0; i < h4; i += th4)
                for (uint32_t j = 0; j < w4; j += tw4) {
                    if (inter && !skip)
                        var_tx(r + i, c + j, max_tx, 0);
                    else
                        stamp_tx(r + i, c + j, tw4, th4, uni);
                }
        }
    }

    // Partitions a decoder meets (decode_partition, §5.11.4): NONE, HORZ,
    // VERT, SPLIT, HORZ_A/B and VERT_A/B from 16x16 up, HORZ_4/VERT_4 for
    // 16x16..64x64; a block crossing the frame edge splits, or halves
    // along the edge it crosses.
    void node_decoder(uint32_t r, uint32_t c, uint32_t s4) {
        if (r >= mi_rows || c >= mi_cols) return;
        if (s4 == 1) {
            leaf(r, c, 1, 1);
            return;
        }
        const uint32_t h = s4 / 2, q = s4 / 4;
        const bool has_rows = r + h < mi_rows, has_cols = c + h < mi_cols;
        auto split = [&] {
            node_decoder(r, c, h);
            node_decoder(r, c + h, h);
            node_decoder(r + h, c, h);
            node_decoder(r + h, c + h, h);
        };
        if (!has_rows || !has_cols) {
            if (has_cols && rng.chance(1, 2)) {  // HORZ: the bottom half is off-frame
                leaf(r, c, s4, h);
            } else if (has_rows && rng.chance(1, 2)) {  // VERT
                leaf(r, c, h, s4);
            } else {
                split();
            }
            return;
        }
        uint32_t split_p = s4 >= 16 ? 45 : (s4 >= 8 ? 35 : 30);
        uint32_t pick = rng.next() % 100;
        if (pick < split_p) {
            split();
            return;
        }
        const bool ab = s4 >= 4, four = s4 >= 4 && s4 <= 16;
        int kind = rng.range(0, 8);  // NONE HORZ VERT HA HB VA VB H4 V4
        if (!ab && kind >= 3) kind = kind % 3;
        if (!four && kind >= 7) kind = kind - 4;
        switch (kind) {
        case 0: leaf(r, c, s4, s4); break;
        case 1: leaf(r, c, s4, h); leaf(r + h, c, s4, h); break;
        case 2: leaf(r, c, h, s4); leaf(r, c + h, h, s4); break;
        case 3: leaf(r, c, h, h); leaf(r, c + h, h, h); leaf(r + h, c, s4, h); break;  // HORZ_A
        case 4: leaf(r, c, s4, h); leaf(r + h, c, h, h); leaf(r + h, c + h, h, h); break;  // HORZ_B
        case 5: leaf(r, c, h, h); leaf(r + h, c, h, h); leaf(r, c + h, h, s4); break;  // VERT_A
        case 6: leaf(r, c, h, s4); leaf(r, c + h, h, h); leaf(r + h, c + h, h, h); break;  // VERT_B
        case 7:  // HORZ_4
            for (uint32_t i = 0; i < 4 && r + i * q < mi_rows; i++) leaf(r + i * q, c, s4, q);
            break;
        default:  // VERT_4
            for (uint32_t i = 0; i < 4 && c + i * q < mi_cols; i++) leaf(r, c + i * q, q, s4);
            break;
        }
    }

    void node(uint32_t r, uint32_t c, uint32_t s4) {
        if (r >= mi_rows || c >= mi_cols) return;
        if (s4 == 1) {
            leaf(r, c, 1, 1);
            return;
        }
        bool outside = r + s4 > mi_rows || c + s4 > mi_cols;
        uint32_t pick = rng.next() % 100;
        // split more often for large blocks and when crossing the frame edge
        uint32_t split_p = s4 >= 16 ? 70 : (s4 >= 8 ? 50 : 40);
        if (outside) split_p = 85;
        uint32_t h = s4 / 2;
        if (pick < split_p) {
            node(r, c, h);
            node(r, c + h, h);
            node(r + h, c, h);
            node(r + h, c + h, h);
        } else if (cfg.rect_partitions && pick < split_p + 10 && s4 <= 32) {
            leaf(r, c, s4, h);  // HORZ
            if (r + h < mi_rows) leaf(r + h, c, s4, h);
        } else if (cfg.rect_partitions && pick < split_p + 20 && s4 <= 32) {
            leaf(r, c, h, s4);  // VERT
            if (c + h < mi_cols) leaf(r, c + h, h, s4);
        } else {
            leaf(r, c, s4, s4);
        }
    }
};

template <typename T>
T *alloc_n(size_t n) {
    T *p = static_cast<T *>(std::calloc(n ? n : 1, sizeof(T)));
    return p;
}

}  // namespace

extern "C" uint64_t oxav1_gen_splitmix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

extern "C" int oxav1_gen_default_config(int cfg, uint64_t seed, oxav1_gen_config *o) {
    if (!o) return OXAV1_ERR_INVALID_ARG;
    std::memset(o, 0, sizeof *o);
    o->bit_depth = 8;
    o->subsampling_x = 1;
    o->subsampling_y = 1;
    o->superres_denom = 8;
    o->enable_lf = o->enable_cdef = o->enable_lr = 1;
    o->noise_amp = 8;
    o->rect_partitions = 1;
    o->lr_types = 7;
    Rng rng(seed ^ 0xC0F1600Dull);
    switch (cfg) {
    case 0:  // 352x288 8-bit 4:2:0, LF+CDEF+LR, 64x64 SB and LR units
        o->upscaled_width = 352;
        o->frame_height = 288;
        break;
    case 1:  // 1080p 8-bit 4:2:0, full chain with film grain (lag 3, 14 points)
        o->upscaled_width = 1920;
        o->frame_height = 1080;
        o->enable_fg = 1;
        o->fg_num_y_points = 14;
        o->fg_ar_lag = 3;
        o->fg_chroma_points = 8;
        break;
    case 2:  // 4K 10-bit 4:2:0, 128 SB, superres 12, LR 128, FG overlap
        o->upscaled_width = 3840;
        o->frame_height = 2160;
        o->bit_depth = 10;
        o->sb128 = 1;
        o->superres_denom = 12;
        o->lr_unit_shift = 1;
        o->lr_types = 6;  // Wiener and SGRPROJ
        o->enable_fg = 1;
        o->fg_overlap = 1;
        o->fg_num_y_points = 14;
        o->fg_ar_lag = 3;
        o->fg_chroma_points = 8;
        o->noise_amp = 32;
        break;
    case 3: {  // batched mix: resolution from the frame index (set by caller)
        static const uint32_t W[4] = {640, 1280, 1920, 3840}, H[4] = {360, 720, 1080, 2160};
        int r = (int)((seed >> 8) & 3);
        o->upscaled_width = W[r];
        o->frame_height = H[r];
        o->bit_depth = rng.chance(1, 2) ? 10 : 8;
        int ss = rng.range(0, 2);
        o->subsampling_x = ss < 2;
        o->subsampling_y = ss == 0;
        o->noise_amp = o->bit_depth == 10 ? 32 : 8;
        break;
    }
    case 4: {  // serving mix: 720p/1080p/4K, 8/10-bit 4:2:0
        static const uint32_t W[3] = {1280, 1920, 3840}, H[3] = {720, 1080, 2160};
        int r = rng.range(0, 2);
        o->upscaled_width = W[r];
        o->frame_height = H[r];
        o->bit_depth = rng.chance(1, 2) ? 10 : 8;
        o->noise_amp = o->bit_depth == 10 ? 32 : 8;
        break;
    }
    default:
        return OXAV1_ERR_INVALID_ARG;
    }
    return OXAV1_OK;
}

extern "C" void oxav1_gen_frame_destroy(oxav1_gen_frame *f) {
    if (!f) return;
    for (int p = 0; p < 3; p++) {
        std::free(f->planes[p]);
        std::free(f->lf_tx_size[p]);
        std::free(f->lr_units[p]);
    }
    std::free(f->is_intra);
    std::free(f->skip);
    std::free(f->ref_frame);
    std::free(f->mode);
    std::free(f->segment_id);
    std::free(f->mi_size);
    std::free(f->delta_lf);
    std::free(f->lf_mi);
    std::free(f->cdef_idx);
    std::free(f);
}

extern "C" int oxav1_gen_frame_create(const oxav1_gen_config *cfg, uint64_t seed,
                                      oxav1_gen_frame **out) {
    if (!cfg || !out) return OXAV1_ERR_INVALID_ARG;
    *out = nullptr;
    if (cfg->upscaled_width == 0 || cfg->frame_height == 0) return OXAV1_ERR_INVALID_ARG;
    if (cfg->bit_depth != 8 && cfg->bit_depth != 10 && cfg->bit_depth != 12)
        return OXAV1_ERR_INVALID_ARG;
    uint32_t denom = cfg->superres_denom ? cfg->superres_denom : 8;
    if (denom < 8 || denom > 16) return OXAV1_ERR_INVALID_ARG;
    auto *f = alloc_n<oxav1_gen_frame>(1);
    if (!f) return OXAV1_ERR_OUT_OF_MEMORY;
    oxav1_frame_desc &d = f->desc;
    const uint32_t sx = cfg->mono ? 1 : cfg->subsampling_x, sy = cfg->mono ? 1 : cfg->subsampling_y;
    d.upscaled_width = cfg->upscaled_width;
    d.frame_height = cfg->frame_height;
    d.frame_width = (cfg->upscaled_width * 8 + denom / 2) / denom;  // frame_header.rs:1742
    d.use_superres = denom != 8 && d.frame_width < d.upscaled_width;
    d.mi_cols = 2 * ((d.frame_width + 7) >> 3);  // frame_header.rs:1745
    d.mi_rows = 2 * ((d.frame_height + 7) >> 3);
    d.bit_depth = cfg->bit_depth;
    d.subsampling_x = (uint8_t)sx;
    d.subsampling_y = (uint8_t)sy;
    d.num_planes = cfg->mono ? 1 : 3;
    d.matrix_coefficients = 1;
    const uint32_t np = d.num_planes, mi_rows = d.mi_rows, mi_cols = d.mi_cols;
    const size_t nmi = (size_t)mi_rows * mi_cols;

    Rng pr(oxav1_gen_splitmix64(seed ^ 0x504C414E45ull));  // planes
    Rng mr(oxav1_gen_splitmix64(seed ^ 0x4D4F444553ull));  // modes
    Rng qr(oxav1_gen_splitmix64(seed ^ 0x504152414Dull));  // params

    // ---- planes: smooth gradient + uniform noise, Clip1 ----
    const int32_t maxv = (1 << d.bit_depth) - 1;
    for (uint32_t p = 0; p < np; p++) {
        uint32_t psx = p ? sx : 0, psy = p ? sy : 0;
        uint32_t rows = (mi_rows * 4) >> psy, cols = (mi_cols * 4) >> psx;
        f->plane_rows[p] = rows;
        f->plane_cols[p] = cols;
        f->planes[p] = alloc_n<int32_t>((size_t)rows * cols);
        if (!f->planes[p]) {
            oxav1_gen_frame_destroy(f);
            return OXAV1_ERR_OUT_OF_MEMORY;
        }
        const int32_t amp = cfg->noise_amp;
        for (uint32_t y = 0; y < rows; y++)
            for (uint32_t x = 0; x < cols; x++) {
                // ((x/W)*0.6 + (y/H)*0.4) * maxv in fixed point, phase-shifted per plane
                int64_t g = ((int64_t)x * 600 / (cols ? cols : 1) + (int64_t)y * 400 / (rows ? rows : 1));
                int32_t base = (int32_t)((g * maxv) / 1000);
                if (p) base = maxv - base / 2 - (int32_t)p * (maxv >> 3);
                int32_t v = base + (amp ? pr.range(-amp, amp) : 0);
                f->planes[p][(size_t)y * cols + x] = v < 0 ? 0 : (v > maxv ? maxv : v);
            }
    }

    // ---- per-mi grids from the partition quadtree ----
    f->is_intra = alloc_n<uint8_t>(nmi);
    f->skip = alloc_n<uint8_t>(nmi);
    f->ref_frame = alloc_n<uint8_t>(nmi);
    f->mode = alloc_n<uint8_t>(nmi);
    f->segment_id = alloc_n<uint8_t>(nmi);
    f->mi_size = alloc_n<uint8_t>(nmi);
    f->delta_lf = alloc_n<int8_t>(nmi * 4);
    f->lf_mi = alloc_n<uint32_t>(nmi);
    if (!f->is_intra || !f->skip || !f->ref_frame || !f->mode || !f->segment_id || !f->mi_size ||
        !f->delta_lf || !f->lf_mi) {
        oxav1_gen_frame_destroy(f);
        return OXAV1_ERR_OUT_OF_MEMORY;
    }
    Builder b(*cfg, mr.next(), f);
    const uint32_t sb4 = cfg->sb128 ? 32 : 16;
    for (uint32_t r = 0; r < mi_rows; r += sb4)
        for (uint32_t c = 0; c < mi_cols; c += sb4) {
            if (cfg->decoder_geometry)
                b.node_decoder(r, c, sb4);
            else
                b.node(r, c, sb4);
        }
    for (size_t k = 0; k < nmi; k++) {
        f->lf_mi[k] = ((uint32_t)f->mi_size[k] << OXAV1_MI_SIZE_SHIFT) |
                      ((uint32_t)(f->skip[k] != 0) << OXAV1_MI_SKIP_BIT) |
                      ((uint32_t)(f->is_intra[k] != 0) << OXAV1_MI_INTRA_BIT) |
                      ((uint32_t)f->ref_frame[k] << OXAV1_MI_REF_SHIFT) |
                      ((uint32_t)f->mode[k] << OXAV1_MI_MODE_SHIFT) |
                      ((uint32_t)f->segment_id[k] << OXAV1_MI_SEG_SHIFT);
    }
    for (uint32_t p = 0; p < 3; p++) {
        uint32_t psx = p ? sx : 0, psy = p ? sy : 0;
        uint32_t tr = (mi_rows + psy) >> psy, tc = (mi_cols + psx) >> psx;
        f->tx_rows[p] = tr;
        f->tx_cols[p] = tc;
        f->lf_tx_size[p] = alloc_n<uint8_t>((size_t)tr * tc);
        if (!f->lf_tx_size[p]) {
            oxav1_gen_frame_destroy(f);
            return OXAV1_ERR_OUT_OF_MEMORY;
        }
        for (uint32_t r = 0; r < tr; r++)
            for (uint32_t c = 0; c < tc; c++) {
                uint32_t lr_ = r << psy, lc = c << psx;
                size_t k = (size_t)lr_ * mi_cols + lc;
                int lt = b.luma_tx[k];
                int v = p == 0 ? lt : chroma_tx((int)p, lt, f->mi_size[k], (int)psx, (int)psy);
                f->lf_tx_size[p][(size_t)r * tc + c] = (uint8_t)(v < 0 ? 0 : v);
            }
    }

    // ---- loop filter params ----
    d.apply_lf = cfg->enable_lf;
    oxav1_lf_ctx &lf = d.lf;
    for (int i = 0; i < 4; i++) lf.loop_filter_level[i] = (uint8_t)qr.range(0, 63);
    if (cfg->enable_lf && lf.loop_filter_level[0] == 0 && lf.loop_filter_level[1] == 0)
        lf.loop_filter_level[0] = (uint8_t)qr.range(1, 63);
    if (np == 1) lf.loop_filter_level[2] = lf.loop_filter_level[3] = 0;
    lf.loop_filter_sharpness = (uint8_t)qr.range(0, 7);
    lf.loop_filter_delta_enabled = qr.chance(1, 2);
    static const int8_t kRefDeltas[8] = {1, 0, 0, 0, -1, 0, -1, -1};  // :239
    std::memcpy(lf.loop_filter_ref_deltas, kRefDeltas, 8);
    lf.loop_filter_mode_deltas[0] = lf.loop_filter_mode_deltas[1] = 0;
    lf.delta_lf_multi = cfg->exercise_delta_lf ? qr.chance(1, 2) : 0;
    lf.num_planes = (uint8_t)np;
    lf.bit_depth = d.bit_depth;
    lf.subsampling_x = (uint8_t)sx;
    lf.subsampling_y = (uint8_t)sy;
    lf.mi_rows = mi_rows;
    lf.mi_cols = mi_cols;
    lf.frame_width = d.frame_width;
    lf.frame_height = d.frame_height;
    if (cfg->exercise_segments) {
        for (int s = 0; s < 8; s++)
            for (int ft = 1; ft <= 4; ft++) {
                d.seg_feature_active[s][ft] = qr.chance(1, 2);
                d.seg_feature_data[s][ft] = (int16_t)qr.range(-63, 63);
            }
    }

    // ---- CDEF ----
    d.apply_cdef = cfg->enable_cdef;
    oxav1_cdef_params &cp = d.cdef;
    static const uint8_t kSec[4] = {0, 1, 2, 4};
    cp.cdef_damping = (uint8_t)qr.range(3, 6);
    cp.cdef_bits = 3;
    for (int i = 0; i < 8; i++) {
        cp.y_pri_strength[i] = (uint8_t)qr.range(0, 15);
        cp.y_sec_strength[i] = kSec[qr.range(0, 3)];
        cp.uv_pri_strength[i] = np > 1 ? (uint8_t)qr.range(0, 15) : 0;
        cp.uv_sec_strength[i] = np > 1 ? kSec[qr.range(0, 3)] : 0;
    }
    f->cdef_rows = (mi_rows + 15) >> 4;
    f->cdef_cols = (mi_cols + 15) >> 4;
    f->cdef_idx = alloc_n<int8_t>((size_t)f->cdef_rows * f->cdef_cols);
    for (size_t i = 0; i < (size_t)f->cdef_rows * f->cdef_cols; i++)
        f->cdef_idx[i] = (cfg->cdef_skip_units && qr.chance(1, 8)) ? (int8_t)-1 : (int8_t)qr.range(0, 7);

    // ---- loop restoration ----
    oxav1_lr_ctx &lr = d.lr;
    lr.mi_rows = mi_rows;
    lr.mi_cols = mi_cols;
    lr.num_planes = (uint8_t)np;
    lr.bit_depth = d.bit_depth;
    lr.subsampling_x = (uint8_t)sx;
    lr.subsampling_y = (uint8_t)sy;
    lr.frame_height = d.frame_height;
    lr.upscaled_width = d.upscaled_width;
    lr.uses_lr = cfg->enable_lr;
    const uint32_t lsize = 64u << (cfg->lr_unit_shift > 2 ? 2 : cfg->lr_unit_shift);
    const uint32_t uvs = (sx && sy) ? (cfg->lr_uv_shift ? 1 : 0) : 0;
    for (uint32_t p = 0; p < 3; p++) {
        lr.frame_restoration_type[p] = (cfg->enable_lr && p < np) ? OXAV1_RESTORE_SWITCHABLE : 0;
        lr.loop_restoration_size[p] = cfg->enable_lr ? (p ? lsize >> uvs : lsize) : 0;
    }
    static const int32_t kWMin[3] = {-5, -23, -17}, kWMax[3] = {10, 8, 46};
    static const int32_t kXMin[2] = {-96, -32}, kXMax[2] = {31, 95};
    for (uint32_t p = 0; p < np && cfg->enable_lr; p++) {
        uint32_t psx = p ? sx : 0, psy = p ? sy : 0;
        uint32_t us = lr.loop_restoration_size[p];
        uint32_t ur = oxav1_pf_count_units_in_frame(us, (d.frame_height + psy) >> psy);
        uint32_t uc = oxav1_pf_count_units_in_frame(us, (d.upscaled_width + psx) >> psx);
        f->lr_unit_rows[p] = ur;
        f->lr_unit_cols[p] = uc;
        f->lr_units[p] = alloc_n<oxav1_lr_unit>((size_t)ur * uc);
        uint8_t allowed[3];
        int na = 0;
        for (int t = 0; t < 3; t++)
            if (cfg->lr_types & (1u << t)) allowed[na++] = (uint8_t)t;
        if (na == 0) allowed[na++] = 0;
        for (size_t i = 0; i < (size_t)ur * uc; i++) {
            oxav1_lr_unit &u = f->lr_units[p][i];
            u.type = allowed[qr.range(0, na - 1)];
            if (u.type == OXAV1_RESTORE_WIENER) {
                for (int ps = 0; ps < 2; ps++)
                    for (int j = 0; j < 3; j++)
                        u.wiener[ps][j] = (p != 0 && j == 0) ? 0 : (int16_t)qr.range(kWMin[j], kWMax[j]);
            } else if (u.type == OXAV1_RESTORE_SGRPROJ) {
                u.sgr_set = (uint8_t)qr.range(0, 15);
                for (int i2 = 0; i2 < 2; i2++) {
                    int32_t radius = oxav1::kSgrParams[u.sgr_set * 4 + i2 * 2];
                    int32_t v;
                    if (radius != 0) {
                        v = qr.range(kXMin[i2], kXMax[i2]);
                    } else if (i2 == 1) {
                        v = 128 - u.sgr_xqd[0];
                        v = v < kXMin[1] ? kXMin[1] : (v > kXMax[1] ? kXMax[1] : v);
                    } else {
                        v = 0;
                    }
                    u.sgr_xqd[i2] = (int16_t)v;
                }
            }
        }
    }

    // ---- film grain ----
    oxav1_film_grain_params &fg = d.fg;
    if (cfg->enable_fg) {
        fg.apply_grain = 1;
        fg.update_grain = 1;
        fg.grain_seed = (uint16_t)qr.next();
        uint32_t ny = cfg->fg_num_y_points > 14 ? 14 : cfg->fg_num_y_points;
        fg.num_y_points = (uint8_t)ny;
        // strictly increasing values spread over 0..255
        uint32_t prev = 0;
        for (uint32_t i = 0; i < ny; i++) {
            uint32_t lo = i == 0 ? 0 : prev + 1;
            uint32_t hi = 255 - (ny - 1 - i) * 2;
            uint32_t span = (hi - lo) / (ny - i);
            uint32_t v = lo + (span ? (uint32_t)qr.range(0, (int32_t)span) : 0);
            fg.point_y_value[i] = (uint8_t)v;
            fg.point_y_scaling[i] = (uint8_t)qr.range(0, 255);
            prev = v;
        }
        fg.chroma_scaling_from_luma = 0;
        uint32_t nc = cfg->fg_chroma_points > 10 ? 10 : cfg->fg_chroma_points;
        if (np == 1 || (sx && sy && ny == 0)) nc = 0;
        fg.num_cb_points = fg.num_cr_points = (uint8_t)nc;
        for (int pl = 0; pl < 2; pl++) {
            uint8_t *vals = pl ? fg.point_cr_value : fg.point_cb_value;
            uint8_t *scal = pl ? fg.point_cr_scaling : fg.point_cb_scaling;
            uint32_t pv = 0;
            for (uint32_t i = 0; i < nc; i++) {
                uint32_t lo = i == 0 ? 0 : pv + 1;
                uint32_t hi = 255 - (nc - 1 - i) * 2;
                uint32_t span = (hi - lo) / (nc - i);
                uint32_t v = lo + (span ? (uint32_t)qr.range(0, (int32_t)span) : 0);
                vals[i] = (uint8_t)v;
                scal[i] = (uint8_t)qr.range(0, 255);
                pv = v;
            }
        }
        fg.grain_scaling = (uint8_t)qr.range(8, 11);
        fg.ar_coeff_lag = cfg->fg_ar_lag > 3 ? 3 : cfg->fg_ar_lag;
        for (int i = 0; i < 24; i++) fg.ar_coeffs_y_plus_128[i] = (uint8_t)qr.range(0, 255);
        for (int i = 0; i < 25; i++) {
            fg.ar_coeffs_cb_plus_128[i] = (uint8_t)qr.range(0, 255);
            fg.ar_coeffs_cr_plus_128[i] = (uint8_t)qr.range(0, 255);
        }
        fg.ar_coeff_shift = (uint8_t)qr.range(6, 9);
        fg.grain_scale_shift = 0;
        fg.cb_mult = (uint8_t)qr.range(0, 255);
        fg.cb_luma_mult = (uint8_t)qr.range(0, 255);
        fg.cb_offset = (uint16_t)qr.range(0, 511);
        fg.cr_mult = (uint8_t)qr.range(0, 255);
        fg.cr_luma_mult = (uint8_t)qr.range(0, 255);
        fg.cr_offset = (uint16_t)qr.range(0, 511);
        fg.overlap_flag = cfg->fg_overlap;
        fg.clip_to_restricted_range = qr.chance(1, 2);
    }
    *out = f;
    return OXAV1_OK;
}
