// Kernels, device job tables and the C ABI of the gfx950 post-filter chain.
//
// Stage kernels take a table of FrameJobs.  Each launch processes work items
// (frame, plane, tile): plain launches cover one frame with one workgroup per
// item; batched launches are persistent (as many workgroups as the kernel's
// occupancy allows on every CU) and pull items from a device counter, so
// frames of any size, bit depth and
// subsampling share one launch per stage.  The batched stage sequence can be
// captured once in a hipGraph and replayed.
//
// The drop-in entry points (oxav1_pf_loop_filter_frame etc.) run the same
// tile functions instantiated for int32 samples on staged copies of the
// caller's PlaneBuffers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <condition_variable>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <unistd.h>
#include <vector>

#include "../../include/oxav1_pf.h"
#include "av1_tables.h"
#include "cdef.hpp"
#include "common.hpp"
#include "copy_pool.hpp"
#include "fg.hpp"
#include "lf.hpp"
#include "lr.hpp"
#include "sr.hpp"

namespace oxav1 {

// ===================================================================== //
// Device side                                                           //
// ===================================================================== //

constexpr int kThreads = 256;

// (frame, local item) for an item of a stage
struct ItemRef {
    uint32_t frame, local;
};
__device__ __forceinline__ ItemRef find_item(const uint32_t *prefix, uint32_t njobs, uint32_t item) {
    if (!prefix) return {0u, item};
    uint32_t lo = 0, hi = njobs;  // prefix[lo] <= item < prefix[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (prefix[mid] <= item)
            lo = mid;
        else
            hi = mid;
    }
    return {lo, item - prefix[lo]};
}
// The job record in the constant address space: its fields become scalar
// loads (s_load through the scalar cache).  Only for reads at 4-byte aligned
// addresses: a byte array indexed by a run-time value would give the scalar
// loads a misaligned base.
using JobC = const __attribute__((address_space(4))) FrameJob;
__device__ __forceinline__ JobC &job_c(const FrameJob &J) { return *(JobC *)&J; }

// plane of item `local` among three planes' item ranges starting at base[0..2]
__device__ __forceinline__ uint32_t plane_of3(const __attribute__((address_space(4))) uint32_t *base, uint32_t local) {
    return local >= base[2] ? 2u : (local >= base[1] ? 1u : 0u);
}
__device__ __forceinline__ uint32_t plane_of(const uint32_t *base, uint32_t nplanes, uint32_t local) {
    uint32_t p = 0;
    while (p + 1 < nplanes && local >= base[p + 1]) p++;
    return p;
}

// --- stage bodies (templated on the sample type) ---
template <typename T>
__device__ __forceinline__ void lf_item(const FrameJob &J, uint32_t local, typename LdsOf<T>::type *region, LfShared &sh) {
    JobC &JK = job_c(J);
    const uint32_t p = plane_of3(JK.item_base[kStageLf], local);
    const uint32_t t = local - JK.item_base[kStageLf][p];
    const uint32_t ty = t / JK.lf_tiles_x[p], tx = t - ty * JK.lf_tiles_x[p];
    lf_tile<T>(J, p, ty, tx, plane_ptr<T>(JK.src[p]), JK.src_pitch[p], plane_ptr<T>(JK.deb[p]), JK.deb_pitch[p], region,
               sh);
}

template <typename T>
__device__ __forceinline__ void cdef_item(const FrameJob &J, uint32_t local, typename LdsOf<T>::type *tile, CdefBlk *blk,
                          CdefTables *tabs) {
    const uint32_t ucols = job_c(J).cdef_ucols;
    const uint32_t ur = local / ucols, uc = local - ur * ucols;
    cdef_unit<T>(J, ur, uc, tile, blk, tabs);
}

// Superres items: planes 0..2 upscale the CDEF frame in runs of
// sr_rows_per_item rows; planes 3..5 upscale the deblocked frame only on the
// four rows around each stripe boundary k >= 1: rows k*(64>>sy) - (8>>sy) - 2
// .. +1 (the rows loop restoration reads from it).
template <typename T>
__device__ __forceinline__ void sr_item(const FrameJob &J, uint32_t local, const int16_t *ftab) {
    const uint32_t p6 = plane_of(J.item_base[kStageSr], 6, local);
    const uint32_t chunk = local - J.item_base[kStageSr][p6];
    const uint32_t p = p6 % 3;
    const bool deb = p6 >= 3;
    int32_t y0, nrows;
    if (!deb) {
        y0 = (int32_t)(chunk * J.sr_rows_per_item);
        nrows = (int32_t)J.sr_rows_per_item;
    } else {
        const int32_t sy = p ? J.sy : 0;
        y0 = (int32_t)(chunk + 1) * (64 >> sy) - (8 >> sy) - 2;
        nrows = 4;
    }
    const G<T> *in = plane_ptr<T>(deb ? (const void *)J.deb[p] : (const void *)J.cdf[p]);
    const uint32_t ip = deb ? J.deb_pitch[p] : J.cdf_pitch[p];
    G<T> *out = plane_ptr<T>(deb ? J.up_deb[p] : J.up_cdef[p]);
    const uint32_t op = deb ? J.up_deb_pitch[p] : J.up_cdef_pitch[p];
    sr_rows<T>(wave_ptr(in), wave_u32(ip), (int32_t)wave_u32(J.pad_cols[p]), wave_ptr(out), wave_u32(op),
               (int32_t)wave_u32(J.out_cols[p]), (int32_t)wave_u32((uint32_t)y0), (int32_t)wave_u32((uint32_t)nrows),
               (int32_t)wave_u32(J.out_rows[p]), (int32_t)wave_u32((uint32_t)J.sr_step[p]),
               (int32_t)wave_u32((uint32_t)J.sr_init[p]), (int32_t)wave_u32((uint32_t)J.sr_maxx[p]), (int32_t)wave_u32(J.bd),
               ftab);
}

template <typename T>
__device__ __forceinline__ void lr_item(const FrameJob &J, uint32_t local, LrSmem<typename LdsOf<T>::type> &S) {
    // the job record is read-only for the whole launch: through the constant
    // address space its fields are scalar loads through the scalar cache
    // instead of vector loads (every base 4-byte aligned: nothing here or in
    // lr_tile reads a byte array indexed by the plane)
    JobC &JK = job_c(J);
    const uint32_t p = plane_of3(JK.item_base[kStageLr], local);
    const uint32_t t = local - JK.item_base[kStageLr][p];
    uint32_t stripe, chunk, tile;
    lr_decode(JK, p, t, stripe, chunk, tile);
    LrStore<T> sink{plane_ptr<T>(JK.ref_out[p]), JK.ref_pitch[p]};
    lr_tile<T, LrStore<T>, JobC>(JK, p, stripe, chunk, tile, S, sink);
}

template <typename T>
__device__ __forceinline__ void fg_item(const FrameJob &J, uint32_t local, uint8_t *lut, const void *&lut_of) {
    // the job record is read-only for the whole launch: through the constant
    // address space its fields are scalar loads (nothing here or in
    // fg_blend_tile indexes a byte array of it by the plane)
    JobC &JK = job_c(J);
    const uint32_t p = plane_of3(JK.item_base[kStageFg], local);
    const uint32_t t = local - JK.item_base[kStageFg][p];
    const uint32_t ty = t / JK.fg_tiles_x[p], tx = t - ty * JK.fg_tiles_x[p];
    // the frame's scaling LUTs, staged again only when the frame changes
    // (jobs without grain have no grain scratch and only copy); the item
    // loop's trailing barrier has ordered the last item's LUT reads, and the
    // condition is uniform in the workgroup, so the barrier after the
    // staging runs only when there was one
    if (JK.apply_fg && JK.fgs != lut_of) {
        for (int i = threadIdx.x; i < 768; i += blockDim.x) lut[i] = ((const G<FgScratch> *)JK.fgs)->lut[i >> 8][i & 255];
        lut_of = JK.fgs;
        __syncthreads();
    }
    const int32_t y0 = (int32_t)ty * 64, x0 = (int32_t)tx * 64;
    const int32_t y1 = min(y0 + 64, (int32_t)JK.out_rows[p]), x1 = min(x0 + 64, (int32_t)JK.out_cols[p]);
    fg_blend_tile<T, JobC>(JK, p, y0, y1, x0, x1, lut);
}

// --- kernels: KIND 0 = engine (u8/u16 by FrameJob::kind), 1 = int32 ---
// counter == nullptr: one workgroup per item (blockIdx.x).  Else persistent:
// claim >= 1 items are taken per atomic on the stage counter (all
// workgroups contend for that one address, so stages with many cheap items
// claim several at once), claim == 0 walks items blockIdx.x + k * gridDim.x
// without atomics.
// A stage's items are split into kClaimParts contiguous partitions, each with
// its own counter on its own 128-byte line: a workgroup claims from the
// partition of its XCD (workgroups are dealt to the XCDs round-robin) and,
// once that is exhausted, from the next ones in turn.  Eight counters take
// the atomics of one; within a partition consecutive claims stay in one
// frame and one XCD's L2.  Sets first/count = (total, 0) when every
// partition is exhausted.
constexpr uint32_t kClaimParts = 8, kClaimStride = 32;
__device__ __forceinline__ void claim_items(uint32_t *counter, uint32_t total, uint32_t claim, uint32_t &part,
                                            uint32_t &first, uint32_t &count) {
    const uint32_t home = blockIdx.x & (kClaimParts - 1);
    for (; part < kClaimParts; part++) {
        const uint32_t r = (home + part) & (kClaimParts - 1);
        const uint32_t lo = (uint32_t)((uint64_t)total * r / kClaimParts);
        const uint32_t hi = (uint32_t)((uint64_t)total * (r + 1) / kClaimParts);
        if (lo >= hi) continue;
        const uint32_t v = lo + atomicAdd(counter + r * kClaimStride, claim);
        if (v < hi) {
            first = v;
            count = min(claim, hi - v);
            return;
        }
    }
    first = total;
    count = 0;
}

// The item loop of every stage kernel.  After thread 0 publishes a claim and
// the workgroup barrier, every thread reads it before the body; the body's
// trailing barrier orders those reads before thread 0 publishes the next one.
#define OXAV1_ITEM_LOOP(BODY)                                                     \
    __shared__ uint32_t s_item, s_n;                                              \
    uint32_t base_ = 0, left_ = 0;                                                \
    uint32_t part_ = 0; /* partitions tried so far (thread 0) */                  \
    uint32_t cf_ = 0, clo_ = 0, chi_ = 0; /* items [clo_, chi_) are frame cf_ */  \
    for (uint32_t iter = 0;; iter++) {                                            \
        uint32_t item;                                                            \
        if (!counter) {                                                           \
            if (iter > 0) break;                                                  \
            item = blockIdx.x;                                                    \
        } else if (claim == 0) {                                                  \
            item = blockIdx.x + iter * gridDim.x;                                 \
        } else {                                                                  \
            if (left_ == 0) {                                                     \
                if (threadIdx.x == 0) claim_items(counter, total, claim, part_, s_item, s_n); \
                __syncthreads();                                                  \
                base_ = __builtin_amdgcn_readfirstlane(s_item);                   \
                left_ = __builtin_amdgcn_readfirstlane(s_n ? s_n : 1u);           \
            }                                                                     \
            item = base_++;                                                       \
            left_--;                                                              \
        }                                                                         \
        if (item >= total) break;                                                 \
        ItemRef ref;                                                              \
        if (!prefix) {                                                            \
            ref = {0u, item};                                                     \
        } else if (item - clo_ < chi_ - clo_) { /* same frame as the last item */ \
            ref = {cf_, item - clo_};                                             \
        } else {                                                                  \
            ref = find_item(prefix, njobs, item);                                 \
            cf_ = ref.frame;                                                      \
            clo_ = item - ref.local;                                              \
            chi_ = prefix[ref.frame + 1];                                         \
        }                                                                         \
        const FrameJob &J = jobs[ref.frame];                                      \
        BODY;                                                                     \
        __syncthreads();                                                          \
    }

template <int KIND>
__device__ __forceinline__ void lf_body(const FrameJob *jobs, const uint32_t *prefix, uint32_t njobs,
                                                 uint32_t total, uint32_t *counter, uint32_t claim) {
    using L = typename std::conditional<KIND == 1, int32_t, uint16_t>::type;
    __shared__ __align__(16) L region[kLfRegion * kLfStride];
    __shared__ LfShared sh;
    OXAV1_ITEM_LOOP({
        if constexpr (KIND == 1) {
            lf_item<int32_t>(J, ref.local, region, sh);
        } else {
            if (job_c(J).kind == kU8)
                lf_item<uint8_t>(J, ref.local, region, sh);
            else
                lf_item<uint16_t>(J, ref.local, region, sh);
        }
    })
}

template <int KIND>
__device__ __forceinline__ void cdef_body(const FrameJob *jobs, const uint32_t *prefix, uint32_t njobs,
                                                 uint32_t total, uint32_t *counter, uint32_t claim) {
    using L = typename std::conditional<KIND == 1, int32_t, uint16_t>::type;
    // u16 tiles are stored twice (packed filter path, kCdefDual)
    __shared__ __align__(16) L tile[KIND == 1 ? kCdefLw * kCdefLw : 2 * kCdefLw * kCdefLw + 2];
    __shared__ CdefBlk blk[64];
    __shared__ CdefTables tabs;
    OXAV1_ITEM_LOOP({
        if constexpr (KIND == 1) {
            cdef_item<int32_t>(J, ref.local, tile, blk, &tabs);
        } else {
            if (job_c(J).kind == kU8)
                cdef_item<uint8_t>(J, ref.local, tile, blk, &tabs);
            else
                cdef_item<uint16_t>(J, ref.local, tile, blk, &tabs);
        }
    })
}

template <int KIND>
__device__ __forceinline__ void sr_body(const FrameJob *jobs, const uint32_t *prefix, uint32_t njobs,
                                                 uint32_t total, uint32_t *counter, uint32_t claim) {
    // the upscale filter table, staged once per workgroup
    __shared__ __align__(16) int16_t ftab[64 * 8];
    for (int i = threadIdx.x; i < 64 * 8; i += blockDim.x) ftab[i] = c_upscale_filter[i >> 3][i & 7];
    __syncthreads();
    OXAV1_ITEM_LOOP({
        if constexpr (KIND == 1) {
            sr_item<int32_t>(J, ref.local, ftab);
        } else {
            if (job_c(J).kind == kU8)
                sr_item<uint8_t>(J, ref.local, ftab);
            else
                sr_item<uint16_t>(J, ref.local, ftab);
        }
    })
}

template <int KIND>
__device__ __forceinline__ void lr_body(const FrameJob *jobs, const uint32_t *prefix, uint32_t njobs,
                                                 uint32_t total, uint32_t *counter, uint32_t claim) {
    using L = typename std::conditional<KIND == 1, int32_t, uint16_t>::type;
    __shared__ LrSmem<L> S;
    OXAV1_ITEM_LOOP({
        if constexpr (KIND == 1) {
            lr_item<int32_t>(J, ref.local, S);
        } else {
            if (job_c(J).kind == kU8)
                lr_item<uint8_t>(J, ref.local, S);
            else
                lr_item<uint16_t>(J, ref.local, S);
        }
    })
}

template <int KIND>
__device__ __forceinline__ void fg_body(const FrameJob *jobs, const uint32_t *prefix, uint32_t njobs,
                                                 uint32_t total, uint32_t *counter, uint32_t claim) {
    __shared__ uint8_t lut[768];
    const void *lut_of = nullptr;  // whose LUTs lut holds (uniform in the workgroup)
    OXAV1_ITEM_LOOP({
        if constexpr (KIND == 1) {
            fg_item<int32_t>(J, ref.local, lut, lut_of);
        } else {
            if (job_c(J).kind == kU8)
                fg_item<uint8_t>(J, ref.local, lut, lut_of);
            else
                fg_item<uint16_t>(J, ref.local, lut, lut_of);
        }
    })
}


// Each stage has a kernel per sample path (KIND) with the compiler's register
// budget, and engine variants k_<stage>_w<W> whose budget allows at least W
// waves per SIMD; the tuning picks one per stage and sample size.
#define OXAV1_STAGE_KERNELS(NAME)                                                                         \
    template <int KIND>                                                                                   \
    __global__ __launch_bounds__(kThreads) void k_##NAME(const FrameJob *jobs, const uint32_t *prefix,   \
                                                         uint32_t njobs, uint32_t total, uint32_t *counter, uint32_t claim) { \
        NAME##_body<KIND>(jobs, prefix, njobs, total, counter, claim);                                          \
    }                                                                                                     \
    template <int W>                                                                                      \
    __global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(W))) void k_##NAME##_w(    \
        const FrameJob *jobs, const uint32_t *prefix, uint32_t njobs, uint32_t total, uint32_t *counter, uint32_t claim) { \
        NAME##_body<0>(jobs, prefix, njobs, total, counter, claim);                                             \
    }
OXAV1_STAGE_KERNELS(lf)
OXAV1_STAGE_KERNELS(cdef)
OXAV1_STAGE_KERNELS(sr)
OXAV1_STAGE_KERNELS(lr)
OXAV1_STAGE_KERNELS(fg)
#undef OXAV1_STAGE_KERNELS

// One workgroup per frame; frames without film grain return at once.
__global__ __launch_bounds__(kThreads) void k_fg_prep(const FrameJob *jobs, uint32_t njobs) {
    __shared__ int16_t sm[4 * kGrainH * kGrainW];  // luma, Cb, Cr templates and the chroma luma term
    if (blockIdx.x >= njobs) return;
    const FrameJob &J = jobs[blockIdx.x];
    if (!J.apply_fg) return;
    fg_prep_frame(J, sm);
}

// ---- side information built on the device from the walker's grids ----
__constant__ uint8_t c_max_tx_rect[22];

// get_tx_size for a chroma plane (cdf.rs:12976-13012): the largest
// rectangular transform of the block's chroma residual size, 64-point
// sides reduced to 32; None (no residual size) maps to TX_4X4 as the
// loop-filter bridge does (cdf.rs:20588-20593).
__device__ __forceinline__ uint32_t dev_chroma_tx(uint32_t mi_size, uint32_t sx, uint32_t sy) {
    if (mi_size >= 22) return 0;
    const uint32_t ps = c_subsampled[mi_size * 4 + sx * 2 + sy];
    if (ps >= 22) return 0;
    const uint32_t uv = c_max_tx_rect[ps];
    const uint32_t w = c_tx_w[uv], h = c_tx_h[uv];
    if (w == 64 || h == 64) return w == 16 ? 9u : (h == 16 ? 10u : 3u);  // 16x32, 32x16, 32x32
    return uv;
}

// One thread per cell: the packed per-mi word and luma transform size of
// every mi (loopfilter_tx_size_at, cdf.rs:20559-20593: inter cells take
// InterTxSizes, intra cells TxSizes), then the chroma transform sizes of
// the co-located luma mi of every chroma cell.
__global__ void k_build_grids(oxav1_walker_grids g, uint32_t mi_rows, uint32_t mi_cols, uint32_t np, uint32_t sx,
                              uint32_t sy, uint32_t *__restrict__ lf_mi, uint8_t *__restrict__ tx0,
                              uint8_t *__restrict__ tx1, uint8_t *__restrict__ tx2) {
    const uint32_t n0 = mi_rows * mi_cols;
    const uint32_t crows = (mi_rows + sy) >> sy, ccols = (mi_cols + sx) >> sx, nc = crows * ccols;
    const uint32_t n = n0 + (np > 1 ? 2 * nc : 0);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (i < n0) {
            const uint32_t inter = g.is_inter[i] != 0;
            lf_mi[i] = ((uint32_t)g.mi_size[i] << OXAV1_MI_SIZE_SHIFT) | ((uint32_t)(g.skip[i] != 0) << OXAV1_MI_SKIP_BIT) |
                       ((inter ^ 1u) << OXAV1_MI_INTRA_BIT) | ((uint32_t)g.ref_frame[i] << OXAV1_MI_REF_SHIFT) |
                       ((uint32_t)g.mode[i] << OXAV1_MI_MODE_SHIFT) | ((uint32_t)g.segment_id[i] << OXAV1_MI_SEG_SHIFT);
            tx0[i] = inter ? g.inter_tx_size[i] : g.tx_size[i];
        } else {
            const uint32_t j = i - n0, plane = 1 + j / nc, k = j - (plane - 1) * nc;
            const uint32_t r = k / ccols, c = k - r * ccols;
            const uint32_t lr = r << sy, lc = c << sx;
            const uint32_t v = lr < mi_rows && lc < mi_cols ? dev_chroma_tx(g.mi_size[lr * mi_cols + lc], sx, sy) : 0u;
            (plane == 1 ? tx1 : tx2)[k] = (uint8_t)v;
        }
    }
}

// Zero the batched work counters (a kernel, so the reset is graph-safe).
// Work counters of one engine: five stages x kClaimParts partitions.
constexpr uint32_t kCounterWords = 5 * kClaimParts * kClaimStride;

__global__ void k_reset(uint32_t *counters, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) counters[i] = 0;
}

// ===================================================================== //
// Host side                                                             //
// ===================================================================== //

thread_local std::string g_last_hip;

static int hip_fail(hipError_t e, const char *what) {
    g_last_hip = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? OXAV1_ERR_OUT_OF_MEMORY : OXAV1_ERR_HIP;
}
#define HIP_TRY(expr)                                    \
    do {                                                 \
        hipError_t e_ = (expr);                          \
        if (e_ != hipSuccess) return hip_fail(e_, #expr); \
    } while (0)

static std::mutex g_init_mu;
static bool g_init_done[64];

// Scratch memory of the allocation-free entry point
// (oxav1_pf_run_frame_noalloc): one frame's intermediates and job record.
// Every device the library runs on holds its own copy until the process
// ends.  Host code never uses the array's name as an address: ensure_device_init
// takes each device's address with hipGetSymbolAddress and keeps it in
// g_scratch_ptr[device].
constexpr size_t kNoAllocScratchBytes = OXAV1_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];
static void *g_scratch_ptr[64];

// Constant tables per device, once.
static int ensure_device_init() {
    int dev = 0;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return OXAV1_ERR_NO_DEVICE;
    HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_init_mu);
    if (dev < 64 && g_init_done[dev]) return OXAV1_OK;
    HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_subsampled), kSubsampledSize, sizeof(kSubsampledSize)));
    HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_max_tx_rect), kMaxTxSizeRect, sizeof(kMaxTxSizeRect)));
    HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_upscale_filter), kUpscaleFilter, sizeof(kUpscaleFilter)));
    HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_sgr_params), kSgrParams, sizeof(kSgrParams)));
    uint16_t a2[256];
    a2[0] = 1;
    a2[255] = 256;
    for (int z = 1; z < 255; z++) a2[z] = (uint16_t)(((z << 8) + z / 2) / (z + 1));
    HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_sgr_a2), a2, sizeof(a2)));
    uint32_t sgr_s[16][2];
    for (int set = 0; set < 16; set++)
        for (int pass = 0; pass < 2; pass++)
            sgr_s[set][pass] = lr_sgr_s_of(kSgrParams[set * 4 + 2 * pass], kSgrParams[set * 4 + 2 * pass + 1]);
    HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_sgr_s), sgr_s, sizeof(sgr_s)));
    uint64_t recip[130];
    for (uint32_t k = 0; k < 130; k++) recip[k] = lr_magic(k);
    HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_lr_recip), recip, sizeof(recip)));
    HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_gauss), kGaussianSequence, sizeof(kGaussianSequence)));
    // LFSR step matrix (film_grain.rs:117-130) and its powers 2^b.
    uint16_t m[16][16];
    for (int i = 0; i < 16; i++) {
        uint32_t col = i >= 1 ? (1u << (i - 1)) : 0u;
        if (i == 0 || i == 1 || i == 3 || i == 12) col |= 1u << 15;
        m[0][i] = (uint16_t)col;
    }
    for (int b = 1; b < 16; b++)
        for (int i = 0; i < 16; i++) {
            // column i of M^(2^b) = M^(2^(b-1)) applied to column i of M^(2^(b-1))
            uint32_t v = m[b - 1][i], r = 0;
            for (int k = 0; k < 16; k++)
                if ((v >> k) & 1u) r ^= m[b - 1][k];
            m[b][i] = (uint16_t)r;
        }
    HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_lfsr_jump), m, sizeof(m)));
    if (dev < 64) {
        void *ptr = nullptr;
        HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = ptr;
        g_init_done[dev] = true;
    }
    return OXAV1_OK;
}

// Grow-only device buffer.
struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
    int ensure(size_t bytes) {
        if (bytes <= n) return OXAV1_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
        if (e != hipSuccess) return hip_fail(e, "hipMalloc");
        n = bytes;
        return OXAV1_OK;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

static inline uint32_t rnd2(uint32_t x, uint32_t n) { return n == 0 ? x : (x + (1u << (n - 1))) >> n; }
static inline uint32_t ceil_div(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

// superres constants (superres.rs:363-382), 64-bit with truncating division.
static void sr_consts(uint32_t dw, uint32_t uw, uint32_t mi_cols, uint32_t sx, int32_t *step, int32_t *init,
                      int32_t *maxx) {
    const int64_t stepx_num = ((int64_t)dw << 14) + (int64_t)(uw / 2);
    const int32_t st = (int32_t)(stepx_num / (int64_t)uw);
    const int32_t err = (int32_t)((int64_t)uw * (int64_t)st - ((int64_t)dw << 14));
    const int64_t diff = (int64_t)uw - (int64_t)dw;
    const int64_t num = -(diff << 13) + (int64_t)(uw / 2);
    const int32_t division = (int32_t)(num / (int64_t)uw);
    int32_t isx = division + (1 << 7) - err / 2;
    isx &= (1 << 14) - 1;
    *step = st;
    *init = isx;
    *maxx = (int32_t)(mi_cols >> sx) * 4 - 1;
}

// LR tile width for a unit size: must divide it; stripes never straddle units.
// pref: 0 = the widest of 64/32 that divides the unit, else 32 or 64 when it divides.
static int lr_tile_width(uint32_t us, uint32_t sy, uint32_t *tw, uint32_t pref = 0) {
    if (us < 32 || us % 32 != 0 || us < (64u >> sy) || us >= 65536u) return OXAV1_ERR_UNSUPPORTED;
    *tw = (us % 64 == 0) ? 64u : 32u;
    if (pref == 32) *tw = 32;
    return OXAV1_OK;
}

static uint32_t lr_stripe_count(uint32_t pey, uint32_t sy) { return ((pey << sy) + 8) / 64 + 1; }

using StageKernel = void (*)(const FrameJob *, const uint32_t *, uint32_t, uint32_t, uint32_t *, uint32_t);
constexpr int kWaveVariants = 4;
constexpr int kWaveOf[kWaveVariants] = {0, 6, 7, 8};  // 0 = the compiler's choice
#define OXAV1_VARIANTS(K) {K<0>, K##_w<6>, K##_w<7>, K##_w<8>}
static StageKernel kEngineKernels[5][kWaveVariants] = {OXAV1_VARIANTS(k_lf), OXAV1_VARIANTS(k_cdef),
                                                       OXAV1_VARIANTS(k_sr), OXAV1_VARIANTS(k_lr),
                                                       OXAV1_VARIANTS(k_fg)};
#undef OXAV1_VARIANTS
static int wave_variant(uint32_t waves) {
    for (int w = 1; w < kWaveVariants; w++)
        if ((uint32_t)kWaveOf[w] == waves) return w;
    return 0;
}

}  // namespace oxav1

using namespace oxav1;

// --------------------------------------------------------------------- //
// Engine                                                                //
// --------------------------------------------------------------------- //

struct FrameSlot {
    DevBuf deb[3], cdf[3], upc[3], upd[3], fgs, offs;
    void release() {
        for (int p = 0; p < 3; p++) {
            deb[p].release();
            cdf[p].release();
            upc[p].release();
            upd[p].release();
        }
        fgs.release();
        offs.release();
    }
};

struct oxav1_pf_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;       // film-grain template branch
    hipEvent_t fork = nullptr, join = nullptr;
    bool own_stream = false;
    uint32_t max_frames = 0;
    int num_cus = 0;
    std::vector<FrameSlot> slots;
    std::vector<FrameJob> jobs_h;
    DevBuf jobs_d, prefix_d, counters_d;
    std::vector<uint32_t> prefix_h;  // 5 x (n+1)
    uint32_t n = 0;
    uint32_t totals[5] = {0, 0, 0, 0, 0};
    bool any_fg = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    bool timing = false;
    hipEvent_t ev[8] = {};
    double stage_sum[6] = {0, 0, 0, 0, 0, 0};
    uint32_t stage_cnt[6] = {0, 0, 0, 0, 0, 0};
    int persistent_blocks[5][kWaveVariants] = {};  // workgroups per CU that fit, per stage and variant
    oxav1_tuning tuning = {};                       // zero = defaults
    int batch_kind = 1;                             // 0 = mostly 8-bit, 1 = mostly 10/12-bit
    // Sub-batches: the batched path split over child engines on their own
    // streams (the children hold job tables built here, no frame slots).
    std::vector<oxav1_pf_engine *> subs;
    std::vector<hipEvent_t> sub_join;
    hipEvent_t sub_fork = nullptr;
    uint32_t nsub = 0;
};

constexpr uint32_t kMaxSubBatches = 4;
constexpr uint8_t kClaimStatic = 255;  // items_per_claim value: no atomics, grid-stride items

static const char *kStageNames[6] = {"lf", "cdef", "sr", "lr", "fg", "fg_prep"};

// Bump allocation of a frame's intermediates inside a fixed region (the
// allocation-free entry point's scratch); 256-byte aligned pieces.
struct ScratchCarve {
    uint8_t *base = nullptr;
    size_t cap = 0, used = 0;
    int take(size_t n, void **out) {
        const size_t off = (used + 255) & ~(size_t)255;
        if (off > cap || n > cap - off) return OXAV1_ERR_SCRATCH_TOO_SMALL;
        *out = base + off;
        used = off + n;
        return OXAV1_OK;
    }
};

// One frame's job record.  Intermediates come from the frame slot (grown on
// demand) or, when carve is given, from the carve's region; then the
// film-grain scratch header is returned in *fg_header for the caller to
// store on the device instead of being copied here.
// Reciprocals of the loop-restoration divisors (lr_mdiv): tile columns and
// unit size per plane.  Item numbers within a plane and unit-row numerators
// stay below 2^24 and the divisors below 2^16 for any frame AV1 can code.
static int lr_set_magic(FrameJob &J) {
    for (uint32_t p = 0; p < 3; p++) {
        if (J.lr_tiles_x[p] >= 65536u || J.lr_unit_size[p] >= 65536u) return OXAV1_ERR_UNSUPPORTED;
        J.lr_magic_tx[p] = lr_magic(J.lr_tiles_x[p]);
        J.lr_magic_us[p] = lr_magic(J.lr_unit_size[p]);
    }
    if (J.item_base[kStageLr][3] >= (1u << 24)) return OXAV1_ERR_UNSUPPORTED;
    // the loop-restoration output addresses rows with 32-bit offsets
    for (uint32_t p = 0; p < 3; p++)
        if ((uint64_t)J.out_rows[p] * J.ref_pitch[p] >= (1ull << 31)) return OXAV1_ERR_UNSUPPORTED;
    return OXAV1_OK;
}

static int build_job(const oxav1_tuning &tun, const oxav1_frame_desc &d, const oxav1_frame_dev &v, FrameSlot &slot,
                     FrameJob &J, ScratchCarve *carve = nullptr, FgScratch *fg_header = nullptr) {
    std::memset(&J, 0, sizeof J);
    auto take = [&](DevBuf &b, size_t n, void **out) -> int {
        if (carve) return carve->take(n, out);
        int st = b.ensure(n);
        if (st) return st;
        *out = b.p;
        return OXAV1_OK;
    };
    if (d.bit_depth != 8 && d.bit_depth != 10 && d.bit_depth != 12) return OXAV1_ERR_INVALID_ARG;
    if (d.num_planes != 1 && d.num_planes != 3) return OXAV1_ERR_INVALID_ARG;
    if (d.mi_rows == 0 || d.mi_cols == 0 || d.frame_width == 0 || d.frame_height == 0) return OXAV1_ERR_INVALID_ARG;
    const uint32_t np = d.num_planes, sx = d.subsampling_x ? 1 : 0, sy = d.subsampling_y ? 1 : 0;
    const size_t B = d.bit_depth == 8 ? 1 : 2;
    J.fw = d.frame_width;
    J.fh = d.frame_height;
    J.uw = d.upscaled_width ? d.upscaled_width : d.frame_width;
    J.mi_rows = d.mi_rows;
    J.mi_cols = d.mi_cols;
    J.bd = d.bit_depth;
    J.sx = (uint8_t)sx;
    J.sy = (uint8_t)sy;
    J.np = (uint8_t)np;
    J.mc = d.matrix_coefficients;
    J.kind = d.bit_depth == 8 ? kU8 : kU16;
    J.use_sr = d.use_superres && J.uw > J.fw;
    J.apply_lf = d.apply_lf;
    J.apply_cdef = d.apply_cdef;
    J.uses_lr = d.lr.uses_lr;
    J.apply_fg = d.fg.apply_grain;
    for (uint32_t p = 0; p < 3; p++) {
        const uint32_t psx = p ? sx : 0, psy = p ? sy : 0;
        J.pad_rows[p] = p < np ? (d.mi_rows * 4) >> psy : 0;
        J.pad_cols[p] = p < np ? (d.mi_cols * 4) >> psx : 0;
        J.out_rows[p] = p < np ? rnd2(J.fh, psy) : 0;
        J.out_cols[p] = p < np ? rnd2(J.uw, psx) : 0;
        if (p < np && (J.out_rows[p] > J.pad_rows[p] || (!J.use_sr && J.out_cols[p] > J.pad_cols[p])))
            return OXAV1_ERR_INVALID_ARG;  // frame_driver.rs:1030-1032
    }
    // ---- deblocking ----
    std::memcpy(J.lf_level, d.lf.loop_filter_level, 4);
    J.lf_sharp = d.lf.loop_filter_sharpness;
    J.lf_delta_enabled = d.lf.loop_filter_delta_enabled;
    J.lf_delta_multi = d.lf.delta_lf_multi;
    std::memcpy(J.lf_ref_deltas, d.lf.loop_filter_ref_deltas, 8);
    std::memcpy(J.lf_mode_deltas, d.lf.loop_filter_mode_deltas, 2);
    std::memcpy(J.seg_active, d.seg_feature_active, sizeof J.seg_active);
    std::memcpy(J.seg_data, d.seg_feature_data, sizeof J.seg_data);
    J.lf_mi = v.lf_mi;
    J.delta_lf = v.delta_lf;
    for (uint32_t p = 0; p < 3; p++) {
        J.tx[p] = v.lf_tx_size[p];
        J.tx_cols[p] = (d.mi_cols + (p ? sx : 0)) >> (p ? sx : 0);
        J.lf_plane_on[p] = J.apply_lf && p < np && (p == 0 || d.lf.loop_filter_level[1 + p] != 0);
    }
    if (J.apply_lf && (!v.lf_mi || !v.lf_tx_size[0] || (np > 1 && (!v.lf_tx_size[1] || !v.lf_tx_size[2]))))
        return OXAV1_ERR_INVALID_ARG;
    // ---- CDEF ----
    J.cdef = d.cdef;
    J.cdef_idx = v.cdef_idx;
    J.cdef_ucols = ceil_div(d.mi_cols, 16);
    J.cdef_urows = ceil_div(d.mi_rows, 16);
    if (J.apply_cdef && (!v.cdef_idx || !v.lf_mi)) return OXAV1_ERR_INVALID_ARG;
    for (uint32_t p = 0; p < np; p++) {
        const uint32_t psx = p ? sx : 0, psy = p ? sy : 0;
        J.cdef_ext_rows[p] = std::min(J.pad_rows[p], (ceil_div(d.mi_rows, 2) * 8) >> psy);
        J.cdef_ext_cols[p] = std::min(J.pad_cols[p], (ceil_div(d.mi_cols, 2) * 8) >> psx);
    }
    // ---- buffers: src -> deb -> cdf -> (SR) -> LR -> ref_out -> FG -> out ----
    for (uint32_t p = 0; p < np; p++) {
        if (!v.src[p] || !v.ref_out[p]) return OXAV1_ERR_INVALID_ARG;
        J.src[p] = v.src[p];
        J.src_pitch[p] = v.src_pitch[p] ? v.src_pitch[p] : J.pad_cols[p];
        const size_t padn = (size_t)J.pad_rows[p] * J.pad_cols[p] * B;
        if (J.lf_plane_on[p]) {
            int st = take(slot.deb[p], padn, &J.deb[p]);
            if (st) return st;
            J.deb_pitch[p] = J.pad_cols[p];
        } else {
            J.deb[p] = const_cast<void *>(v.src[p]);
            J.deb_pitch[p] = J.src_pitch[p];
        }
        if (J.apply_cdef) {
            int st = take(slot.cdf[p], padn, &J.cdf[p]);
            if (st) return st;
            J.cdf_pitch[p] = J.pad_cols[p];
        } else {
            J.cdf[p] = J.deb[p];
            J.cdf_pitch[p] = J.deb_pitch[p];
        }
    }
    // ---- superres ----
    {
        const uint32_t r = tun.sr_rows_per_item[J.kind == kU8 ? 0 : 1];
        J.sr_rows_per_item = r ? r : 8;
    }
    if (J.use_sr) {
        for (uint32_t p = 0; p < np; p++) {
            const uint32_t psx = p ? sx : 0;
            const uint32_t dw = rnd2(J.fw, psx), uw = rnd2(J.uw, psx);
            if (uw <= dw) return OXAV1_ERR_UPSCALED_NOT_LARGER;
            sr_consts(dw, uw, d.mi_cols, psx, &J.sr_step[p], &J.sr_init[p], &J.sr_maxx[p]);
            J.sr_in_w[p] = dw;
            const size_t outn = (size_t)J.out_rows[p] * J.out_cols[p] * B;
            int st = take(slot.upc[p], outn, &J.up_cdef[p]);
            if (st) return st;
            J.up_cdef_pitch[p] = J.out_cols[p];
            if (J.uses_lr) {
                st = take(slot.upd[p], outn, &J.up_deb[p]);
                if (st) return st;
                J.up_deb_pitch[p] = J.out_cols[p];
            }
            J.lr_rows[p] = J.lrc_rows[p] = J.out_rows[p];
            J.lr_cols[p] = J.lrc_cols[p] = J.out_cols[p];
        }
        J.sr_deb_on = J.uses_lr;
    } else {
        for (uint32_t p = 0; p < np; p++) {
            J.up_cdef[p] = J.cdf[p];
            J.up_cdef_pitch[p] = J.cdf_pitch[p];
            J.up_deb[p] = J.deb[p];
            J.up_deb_pitch[p] = J.deb_pitch[p];
            J.lr_rows[p] = J.lrc_rows[p] = J.pad_rows[p];
            J.lr_cols[p] = J.lrc_cols[p] = J.pad_cols[p];
        }
    }
    // ---- loop restoration ----
    // A job whose pre-grain frame is already final (no stage before the grain
    // runs and the caller passes the stored frame as both src and ref_out:
    // the show_existing_frame output, frame_driver.rs:1489-1537) skips the
    // copy into ref_out.
    bool in_place = !J.apply_lf && !J.apply_cdef && !J.use_sr && !J.uses_lr;
    for (uint32_t p = 0; p < np; p++)
        in_place = in_place && v.src[p] == v.ref_out[p] &&
                   J.src_pitch[p] == (v.ref_pitch[p] ? v.ref_pitch[p] : J.out_cols[p]);
    J.lr_copy_none = in_place ? 0 : 1;
    for (uint32_t p = 0; p < np; p++) {
        const uint32_t psx = p ? sx : 0, psy = p ? sy : 0;
        J.lr_type[p] = J.uses_lr ? d.lr.frame_restoration_type[p] : OXAV1_RESTORE_NONE;
        J.lr_units[p] = v.lr_units[p];
        J.lr_tile_w[p] = 64;
        if (J.lr_type[p] != OXAV1_RESTORE_NONE) {
            const uint32_t us = d.lr.loop_restoration_size[p];
            int st = lr_tile_width(us, psy, &J.lr_tile_w[p], tun.lr_tile_w[J.kind == kU8 ? 0 : 1]);
            if (st) return st;
            J.lr_unit_size[p] = us;
            J.lr_unit_rows[p] = oxav1_pf_count_units_in_frame(us, rnd2(J.fh, psy));
            J.lr_unit_cols[p] = oxav1_pf_count_units_in_frame(us, rnd2(J.uw, psx));
        }
        J.ref_out[p] = v.ref_out[p];
        J.ref_pitch[p] = v.ref_pitch[p] ? v.ref_pitch[p] : J.out_cols[p];
    }
    // ---- film grain ----
    for (uint32_t p = 0; p < np; p++) {
        J.out[p] = v.out[p] ? v.out[p] : v.ref_out[p];
        J.out_pitch[p] = v.out[p] ? (v.out_pitch[p] ? v.out_pitch[p] : J.out_cols[p]) : J.ref_pitch[p];
        if (J.apply_fg && J.out[p] == J.ref_out[p]) return OXAV1_ERR_ALIAS;
    }
    J.fg = d.fg;
    const bool fg_stage = J.apply_fg || (J.out[0] != J.ref_out[0]);
    if (J.apply_fg) {
        void *fgs = nullptr, *offs = nullptr;
        int st = take(slot.fgs, sizeof(FgScratch), &fgs);
        if (st) return st;
        const uint32_t nst = ceil_div(ceil_div(J.fh, 2), 16), nbl = ceil_div(ceil_div(J.uw, 2), 16);
        st = take(slot.offs, (size_t)nst * nbl, &offs);
        if (st) return st;
        J.fgs = (FgScratch *)fgs;
        FgScratch hs;
        std::memset(&hs, 0, sizeof hs);
        hs.nstripes = nst;
        hs.nblocks = nbl;
        hs.offsets = (uint8_t *)offs;
        if (carve) {
            if (fg_header) *fg_header = hs;
        } else {
            HIP_TRY(hipMemcpy(fgs, &hs, sizeof hs, hipMemcpyHostToDevice));
        }
    }
    // ---- item tables ----
    uint32_t acc = 0;
    for (uint32_t p = 0; p < 3; p++) {
        J.item_base[kStageLf][p] = acc;
        J.lf_tiles_x[p] = ceil_div(J.pad_cols[p], kLfTile);
        if (p < np && J.lf_plane_on[p]) acc += ceil_div(J.pad_rows[p], kLfTile) * J.lf_tiles_x[p];
    }
    J.item_base[kStageLf][3] = acc;
    J.item_base[kStageCdef][0] = 0;
    J.item_base[kStageCdef][1] = J.apply_cdef ? J.cdef_urows * J.cdef_ucols : 0;
    acc = 0;
    for (uint32_t p6 = 0; p6 < 6; p6++) {
        J.item_base[kStageSr][p6] = acc;
        const uint32_t p = p6 % 3;
        const bool on = J.use_sr && p < np && (p6 < 3 || J.sr_deb_on);
        if (on && p6 < 3) acc += ceil_div(J.out_rows[p], J.sr_rows_per_item);
        if (on && p6 >= 3) {  // stripe-boundary row groups below the last output row
            const uint32_t sy = p ? J.sy : 0;
            acc += (J.out_rows[p] - 1 + (8u >> sy) + 2) / (64u >> sy);
        }
    }
    J.item_base[kStageSr][6] = acc;
    acc = 0;
    for (uint32_t p = 0; p < 3; p++) {
        J.item_base[kStageLr][p] = acc;
        if (p < np) {
            const uint32_t psy = p ? sy : 0;
            J.lr_tiles_x[p] = ceil_div(J.out_cols[p], J.lr_tile_w[p]);
            if (J.uses_lr || J.lr_copy_none)
                acc += lr_stripe_count(J.out_rows[p] - 1, psy) * lr_chunks(psy) * J.lr_tiles_x[p];
        }
    }
    J.item_base[kStageLr][3] = acc;
    if (int st = lr_set_magic(J)) return st;
    acc = 0;
    for (uint32_t p = 0; p < 3; p++) {
        J.item_base[kStageFg][p] = acc;
        J.fg_tiles_x[p] = ceil_div(J.out_cols[p], 64);
        if (p < np && fg_stage) acc += ceil_div(J.out_rows[p], 64) * J.fg_tiles_x[p];
    }
    J.item_base[kStageFg][3] = acc;
    return OXAV1_OK;
}

static uint32_t stage_items(const FrameJob &J, int s) {
    switch (s) {
    case kStageLf: return J.item_base[kStageLf][3];
    case kStageCdef: return J.item_base[kStageCdef][1];
    case kStageSr: return J.item_base[kStageSr][6];
    case kStageLr: return J.item_base[kStageLr][3];
    default: return J.item_base[kStageFg][3];
    }
}

extern "C" int oxav1_pf_engine_create(oxav1_pf_engine **out, void *hip_stream, uint32_t max_frames) {
    if (!out) return OXAV1_ERR_INVALID_ARG;
    *out = nullptr;
    int st = ensure_device_init();
    if (st) return st;
    auto *e = new (std::nothrow) oxav1_pf_engine();
    if (!e) return OXAV1_ERR_OUT_OF_MEMORY;
    (void)hipGetDevice(&e->device);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, e->device) == hipSuccess) e->num_cus = prop.multiProcessorCount;
    if (e->num_cus <= 0) e->num_cus = 256;
    // persistent grid per stage: every workgroup slot the kernel can occupy
    {
        for (int s = 0; s < 5; s++)
            for (int w = 0; w < kWaveVariants; w++) {
                int per_cu = 0;
                if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)kEngineKernels[s][w], kThreads,
                                                                 0) != hipSuccess ||
                    per_cu <= 0)
                    per_cu = 2;
                e->persistent_blocks[s][w] = per_cu;
            }
    }
    e->max_frames = max_frames ? max_frames : 1;
    if (hip_stream) {
        e->stream = (hipStream_t)hip_stream;
    } else {
        if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
            delete e;
            return OXAV1_ERR_HIP;
        }
        e->own_stream = true;
    }
    for (auto &ev : e->ev)
        if (hipEventCreate(&ev) != hipSuccess) {
            delete e;
            return OXAV1_ERR_HIP;
        }
    if (hipStreamCreateWithFlags(&e->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&e->join, hipEventDisableTiming) != hipSuccess) {
        delete e;
        return OXAV1_ERR_HIP;
    }
    st = e->counters_d.ensure(kCounterWords * sizeof(uint32_t));
    if (st) {
        delete e;
        return st;
    }
    *out = e;
    return OXAV1_OK;
}

extern "C" int oxav1_pf_engine_destroy(oxav1_pf_engine *e) {
    if (!e) return OXAV1_OK;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    for (auto *c : e->subs) (void)oxav1_pf_engine_destroy(c);
    for (auto ev : e->sub_join)
        if (ev) (void)hipEventDestroy(ev);
    if (e->sub_fork) (void)hipEventDestroy(e->sub_fork);
    if (e->graph_exec) (void)hipGraphExecDestroy(e->graph_exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    for (auto &s : e->slots) s.release();
    e->jobs_d.release();
    e->prefix_d.release();
    e->counters_d.release();
    for (auto &ev : e->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (e->fork) (void)hipEventDestroy(e->fork);
    if (e->join) (void)hipEventDestroy(e->join);
    if (e->side) (void)hipStreamDestroy(e->side);
    if (e->own_stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return OXAV1_OK;
}

static int load_batch(oxav1_pf_engine *e);
static int split_batch(oxav1_pf_engine *e);

extern "C" int oxav1_pf_engine_prepare(oxav1_pf_engine *e, const oxav1_frame_desc *desc, const oxav1_frame_dev *dev,
                                       uint32_t n) {
    if (!e || !desc || !dev || n == 0) return OXAV1_ERR_INVALID_ARG;
    if (n > e->max_frames) return OXAV1_ERR_ADMISSION_REFUSED;
    HIP_TRY(hipSetDevice(e->device));
    HIP_TRY(hipStreamSynchronize(e->stream));
    if (e->slots.size() < n) e->slots.resize(n);
    e->jobs_h.resize(n);
    e->nsub = 0;
    for (uint32_t f = 0; f < n; f++) {
        int st = build_job(e->tuning, desc[f], dev[f], e->slots[f], e->jobs_h[f]);
        if (st) return st;
    }
    int st = load_batch(e);
    if (st) return st;
    return split_batch(e);
}

// Device job table and per-stage prefix tables of the batch in e->jobs_h.
static int load_batch(oxav1_pf_engine *e) {
    const uint32_t n = (uint32_t)e->jobs_h.size();
    e->n = n;
    e->any_fg = false;
    for (const FrameJob &J : e->jobs_h) e->any_fg |= J.apply_fg != 0;
    {
        uint64_t px[2] = {0, 0};
        for (uint32_t f = 0; f < n; f++)
            px[e->jobs_h[f].kind == kU8 ? 0 : 1] += (uint64_t)e->jobs_h[f].uw * e->jobs_h[f].fh;
        e->batch_kind = px[0] > px[1] ? 0 : 1;
    }
    e->prefix_h.assign(5 * (n + 1), 0);
    for (int s = 0; s < 5; s++) {
        uint32_t acc = 0;
        for (uint32_t f = 0; f < n; f++) {
            e->prefix_h[s * (n + 1) + f] = acc;
            acc += stage_items(e->jobs_h[f], s);
        }
        e->prefix_h[s * (n + 1) + n] = acc;
        e->totals[s] = acc;
    }
    int st = e->jobs_d.ensure(sizeof(FrameJob) * n);
    if (st) return st;
    st = e->prefix_d.ensure(sizeof(uint32_t) * e->prefix_h.size());
    if (st) return st;
    HIP_TRY(hipMemcpy(e->jobs_d.p, e->jobs_h.data(), sizeof(FrameJob) * n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(e->prefix_d.p, e->prefix_h.data(), sizeof(uint32_t) * e->prefix_h.size(),
                      hipMemcpyHostToDevice));
    if (e->graph_exec) (void)hipGraphExecDestroy(e->graph_exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->graph_exec = nullptr;
    e->graph = nullptr;
    return OXAV1_OK;
}

// Split the prepared batch into sub-batches of balanced output pixels, one
// per child engine, when the tuning asks for more than one: stages of
// different frames then run concurrently on the children's streams, which
// fills the issue slots one stage leaves idle (tails of the persistent
// launches, latency-bound stages beside issue-bound ones).
static int split_batch(oxav1_pf_engine *e) {
    e->nsub = 0;
    const uint32_t n = e->n;
    const uint32_t want = std::min<uint32_t>(e->tuning.sub_batches[e->batch_kind], kMaxSubBatches);
    if (want < 2 || n < 2) return OXAV1_OK;
    const uint32_t K = std::min(want, n);
    std::vector<uint32_t> order(n);
    for (uint32_t f = 0; f < n; f++) order[f] = f;
    auto px = [&](uint32_t f) { return (uint64_t)e->jobs_h[f].uw * e->jobs_h[f].fh * e->jobs_h[f].np; };
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return px(a) > px(b); });
    std::vector<std::vector<uint32_t>> grp(K);
    std::vector<uint64_t> load(K, 0);
    for (uint32_t f : order) {
        uint32_t g = 0;
        for (uint32_t k = 1; k < K; k++)
            if (load[k] < load[g]) g = k;
        grp[g].push_back(f);
        load[g] += px(f);
    }
    if (!e->sub_fork && hipEventCreateWithFlags(&e->sub_fork, hipEventDisableTiming) != hipSuccess)
        return OXAV1_ERR_HIP;
    while (e->subs.size() < K) {
        oxav1_pf_engine *c = nullptr;
        int st = oxav1_pf_engine_create(&c, nullptr, e->max_frames);
        if (st) return st;
        e->subs.push_back(c);
        hipEvent_t ev = nullptr;
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) return OXAV1_ERR_HIP;
        e->sub_join.push_back(ev);
    }
    for (uint32_t k = 0; k < K; k++) {
        oxav1_pf_engine *c = e->subs[k];
        HIP_TRY(hipStreamSynchronize(c->stream));
        c->tuning = e->tuning;
        c->tuning.sub_batches[0] = c->tuning.sub_batches[1] = 0;
        c->jobs_h.clear();
        for (uint32_t f : grp[k]) c->jobs_h.push_back(e->jobs_h[f]);
        int st = load_batch(c);
        if (st) return st;
    }
    e->nsub = K;
    return OXAV1_OK;
}


// Plain path: per frame, one launch per stage sized to the frame.
static int enqueue_plain(oxav1_pf_engine *e, bool timed) {
    const FrameJob *jd = (const FrameJob *)e->jobs_d.p;
    for (uint32_t f = 0; f < e->n; f++) {
        const FrameJob &J = e->jobs_h[f];
        if (J.apply_fg) {  // grain templates on the side branch, joined before the blend
            HIP_TRY(hipEventRecord(e->fork, e->stream));
            HIP_TRY(hipStreamWaitEvent(e->side, e->fork, 0));
            hipLaunchKernelGGL(k_fg_prep, dim3(1), dim3(kThreads), 0, e->side, jd + f, 1u);
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipEventRecord(e->join, e->side));
        }
        if (timed && f == 0) HIP_TRY(hipEventRecord(e->ev[0], e->stream));
        for (int s = 0; s < 5; s++) {
            if (s == kStageFg && J.apply_fg) HIP_TRY(hipStreamWaitEvent(e->stream, e->join, 0));
            const uint32_t items = stage_items(J, s);
            if (items) {
                const int w = wave_variant(e->tuning.waves_per_eu[s][J.kind == kU8 ? 0 : 1]);
                hipLaunchKernelGGL(kEngineKernels[s][w], dim3(items), dim3(kThreads), 0, e->stream, jd + f,
                                   (const uint32_t *)nullptr, 1u, items, (uint32_t *)nullptr, 1u);
                HIP_TRY(hipGetLastError());
            }
            if (timed && f == 0) HIP_TRY(hipEventRecord(e->ev[s + 1], e->stream));
        }
    }
    return OXAV1_OK;
}

// Batched path: one persistent launch per stage over every frame.
static int enqueue_batch(oxav1_pf_engine *e, bool timed) {
    const FrameJob *jd = (const FrameJob *)e->jobs_d.p;
    uint32_t *ctr = (uint32_t *)e->counters_d.p;
    const uint32_t *pre = (const uint32_t *)e->prefix_d.p;
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(256), 0, e->stream, ctr, kCounterWords);
    HIP_TRY(hipGetLastError());
    // Grain templates depend only on the parameters: build them on a side
    // branch that overlaps deblock..restoration, joined before the blend.
    if (e->any_fg) {
        HIP_TRY(hipEventRecord(e->fork, e->stream));
        HIP_TRY(hipStreamWaitEvent(e->side, e->fork, 0));
        hipLaunchKernelGGL(k_fg_prep, dim3(e->n), dim3(kThreads), 0, e->side, jd, e->n);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(e->join, e->side));
    }
    if (timed) HIP_TRY(hipEventRecord(e->ev[0], e->stream));
    for (int s = 0; s < 5; s++) {
        const uint32_t total = e->totals[s];
        if (s == kStageFg && e->any_fg) HIP_TRY(hipStreamWaitEvent(e->stream, e->join, 0));
        if (total) {
            const int w = wave_variant(e->tuning.waves_per_eu[s][e->batch_kind]);
            const uint8_t cl = e->tuning.items_per_claim[s][e->batch_kind];
            const uint32_t claim = cl == 0 ? 1u : (cl == kClaimStatic ? 0u : cl);
            const uint32_t per_cu = e->tuning.blocks_per_cu[s][e->batch_kind];
            const uint32_t fit = (uint32_t)e->persistent_blocks[s][w];
            const uint32_t grid =
                std::min<uint32_t>(total, (uint32_t)e->num_cus * (per_cu ? std::min(per_cu, fit) : fit));
            hipLaunchKernelGGL(kEngineKernels[s][w], dim3(grid), dim3(kThreads), 0, e->stream, jd, pre + s * (e->n + 1),
                               e->n, total, ctr + s * kClaimParts * kClaimStride, claim);
            HIP_TRY(hipGetLastError());
        }
        if (timed) HIP_TRY(hipEventRecord(e->ev[s + 1], e->stream));
    }
    return OXAV1_OK;
}

static int collect_timing(oxav1_pf_engine *e) {
    HIP_TRY(hipEventSynchronize(e->ev[5]));
    for (int s = 0; s < 5; s++) {
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, e->ev[s], e->ev[s + 1]));
        e->stage_sum[s] += ms;
        e->stage_cnt[s]++;
    }
    return OXAV1_OK;
}

// mode 0 = plain, 1 = batched, 2 = batched replayed from a hipGraph
extern "C" int oxav1_pf_engine_run(oxav1_pf_engine *e, int mode) {
    if (!e || e->n == 0) return OXAV1_ERR_INVALID_ARG;
    HIP_TRY(hipSetDevice(e->device));
    if (mode == 0) {
        int st = enqueue_plain(e, e->timing);
        if (st) return st;
        return e->timing ? collect_timing(e) : OXAV1_OK;
    }
    if ((mode == 1 || mode == 2) && !e->timing && e->nsub > 1) {
        HIP_TRY(hipEventRecord(e->sub_fork, e->stream));
        for (uint32_t k = 0; k < e->nsub; k++) {
            oxav1_pf_engine *c = e->subs[k];
            HIP_TRY(hipStreamWaitEvent(c->stream, e->sub_fork, 0));
            int st = oxav1_pf_engine_run(c, mode);
            if (st) return st;
            HIP_TRY(hipEventRecord(e->sub_join[k], c->stream));
            HIP_TRY(hipStreamWaitEvent(e->stream, e->sub_join[k], 0));
        }
        return OXAV1_OK;
    }
    if (mode == 1 || e->timing) {
        int st = enqueue_batch(e, e->timing);
        if (st) return st;
        return e->timing ? collect_timing(e) : OXAV1_OK;
    }
    if (mode != 2) return OXAV1_ERR_INVALID_ARG;
    if (!e->graph_exec) {
        HIP_TRY(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
        int st = enqueue_batch(e, false);
        hipGraph_t g = nullptr;
        hipError_t ce = hipStreamEndCapture(e->stream, &g);
        if (st) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        if (ce != hipSuccess) return hip_fail(ce, "hipStreamEndCapture");
        e->graph = g;
        HIP_TRY(hipGraphInstantiate(&e->graph_exec, e->graph, nullptr, nullptr, 0));
    }
    HIP_TRY(hipGraphLaunch(e->graph_exec, e->stream));
    return OXAV1_OK;
}

extern "C" int oxav1_pf_engine_run_frame(oxav1_pf_engine *e, const oxav1_frame_desc *desc,
                                         const oxav1_frame_dev *dev) {
    int st = oxav1_pf_engine_prepare(e, desc, dev, 1);
    return st ? st : oxav1_pf_engine_run(e, 0);
}

extern "C" int oxav1_pf_engine_run_batch(oxav1_pf_engine *e, const oxav1_frame_desc *desc,
                                         const oxav1_frame_dev *dev, uint32_t n, int use_graph) {
    int st = oxav1_pf_engine_prepare(e, desc, dev, n);
    return st ? st : oxav1_pf_engine_run(e, use_graph ? 2 : 1);
}

extern "C" int oxav1_pf_engine_sync(oxav1_pf_engine *e) {
    if (!e) return OXAV1_ERR_INVALID_ARG;
    HIP_TRY(hipStreamSynchronize(e->stream));
    return OXAV1_OK;
}

extern "C" int oxav1_pf_engine_set_timing(oxav1_pf_engine *e, int enable) {
    if (!e) return OXAV1_ERR_INVALID_ARG;
    e->timing = enable != 0;
    for (int s = 0; s < 6; s++) {
        e->stage_sum[s] = 0;
        e->stage_cnt[s] = 0;
    }
    return OXAV1_OK;
}

extern "C" int oxav1_pf_engine_stage_ms(oxav1_pf_engine *e, const char *stage, double *avg_ms, uint32_t *count) {
    if (!e || !stage || !avg_ms) return OXAV1_ERR_INVALID_ARG;
    for (int s = 0; s < 5; s++)
        if (!std::strcmp(stage, kStageNames[s])) {
            *avg_ms = e->stage_cnt[s] ? e->stage_sum[s] / e->stage_cnt[s] : 0.0;
            if (count) *count = e->stage_cnt[s];
            return OXAV1_OK;
        }
    return OXAV1_ERR_INVALID_ARG;
}

extern "C" void *oxav1_pf_engine_stream(oxav1_pf_engine *e) { return e ? (void *)e->stream : nullptr; }

// Size n frame slots for frames of desc's shape with every stage on, so that
// later prepare calls with frames no larger never allocate (a grow would
// hipFree, which waits for the whole device).
extern "C" int oxav1_pf_engine_reserve(oxav1_pf_engine *e, const oxav1_frame_desc *desc, uint32_t n,
                                       uint64_t *bytes) {
    if (!e || !desc || n == 0 || n > e->max_frames) return OXAV1_ERR_INVALID_ARG;
    const oxav1_frame_desc &d = *desc;
    if ((d.bit_depth != 8 && d.bit_depth != 10 && d.bit_depth != 12) || (d.num_planes != 1 && d.num_planes != 3) ||
        d.mi_rows == 0 || d.mi_cols == 0 || d.frame_height == 0)
        return OXAV1_ERR_INVALID_ARG;
    HIP_TRY(hipSetDevice(e->device));
    HIP_TRY(hipStreamSynchronize(e->stream));
    if (e->slots.size() < n) e->slots.resize(n);
    const uint32_t sx = d.subsampling_x ? 1 : 0, sy = d.subsampling_y ? 1 : 0;
    const size_t B = d.bit_depth == 8 ? 1 : 2;
    const uint32_t uw = std::max(d.upscaled_width, d.frame_width);
    const bool sr = d.use_superres && uw > d.frame_width;
    uint64_t total = 0;
    for (uint32_t f = 0; f < n; f++) {
        FrameSlot &slot = e->slots[f];
        for (uint32_t p = 0; p < d.num_planes; p++) {
            const uint32_t psx = p ? sx : 0, psy = p ? sy : 0;
            const size_t padn = (size_t)((d.mi_rows * 4) >> psy) * ((d.mi_cols * 4) >> psx) * B;
            const size_t outn = (size_t)rnd2(d.frame_height, psy) * rnd2(uw, psx) * B;
            int st = slot.deb[p].ensure(padn);
            if (!st) st = slot.cdf[p].ensure(padn);
            if (!st && sr) st = slot.upc[p].ensure(outn);
            if (!st && sr) st = slot.upd[p].ensure(outn);
            if (st) return st;
        }
        const uint32_t nst = ceil_div(ceil_div(d.frame_height, 2), 16), nbl = ceil_div(ceil_div(uw, 2), 16);
        int st = slot.fgs.ensure(sizeof(FgScratch));
        if (!st) st = slot.offs.ensure((size_t)nst * nbl);
        if (st) return st;
        for (uint32_t p = 0; p < 3; p++) total += slot.deb[p].n + slot.cdf[p].n + slot.upc[p].n + slot.upd[p].n;
        total += slot.fgs.n + slot.offs.n;
    }
    if (bytes) *bytes = total;
    return OXAV1_OK;
}

extern "C" int oxav1_pf_engine_set_tuning(oxav1_pf_engine *e, const oxav1_tuning *t) {
    if (!e || !t) return OXAV1_ERR_INVALID_ARG;
    for (int k = 0; k < 2; k++) {
        for (int st = 0; st < 5; st++)
            if (t->waves_per_eu[st][k] && !wave_variant(t->waves_per_eu[st][k])) return OXAV1_ERR_INVALID_ARG;
        if (t->sr_rows_per_item[k] > 64 || t->sub_batches[k] > kMaxSubBatches) return OXAV1_ERR_INVALID_ARG;
        for (int st = 0; st < 5; st++)
            if (t->items_per_claim[st][k] > 16 && t->items_per_claim[st][k] != kClaimStatic)
                return OXAV1_ERR_INVALID_ARG;
        if (t->lr_tile_w[k] != 0 && t->lr_tile_w[k] != 32 && t->lr_tile_w[k] != 64) return OXAV1_ERR_INVALID_ARG;
    }
    e->tuning = *t;
    return OXAV1_OK;
}

extern "C" int oxav1_pf_engine_get_tuning(oxav1_pf_engine *e, oxav1_tuning *t, uint32_t *num_cus,
                                          uint32_t fit_blocks_per_cu[5][4]) {
    if (!e) return OXAV1_ERR_INVALID_ARG;
    if (t) *t = e->tuning;
    if (num_cus) *num_cus = (uint32_t)e->num_cus;
    if (fit_blocks_per_cu)
        for (int s = 0; s < 5; s++)
            for (int w = 0; w < kWaveVariants; w++) fit_blocks_per_cu[s][w] = (uint32_t)e->persistent_blocks[s][w];
    return OXAV1_OK;
}

extern "C" int oxav1_pf_build_grids(const oxav1_walker_grids *g, uint32_t mi_rows, uint32_t mi_cols,
                                    uint8_t num_planes, uint8_t subsampling_x, uint8_t subsampling_y, uint32_t *lf_mi,
                                    uint8_t *const lf_tx_size[3], void *hip_stream) {
    if (!g || !lf_mi || !lf_tx_size || !lf_tx_size[0] || mi_rows == 0 || mi_cols == 0 ||
        (num_planes != 1 && num_planes != 3))
        return OXAV1_ERR_INVALID_ARG;
    if (!g->is_inter || !g->skip || !g->ref_frame || !g->mode || !g->segment_id || !g->mi_size || !g->tx_size ||
        !g->inter_tx_size || (num_planes > 1 && (!lf_tx_size[1] || !lf_tx_size[2])))
        return OXAV1_ERR_INVALID_ARG;
    int st = ensure_device_init();
    if (st) return st;
    const uint32_t sx = subsampling_x ? 1 : 0, sy = subsampling_y ? 1 : 0;
    const uint64_t n = (uint64_t)mi_rows * mi_cols +
                       (num_planes > 1 ? 2ull * ((mi_rows + sy) >> sy) * ((mi_cols + sx) >> sx) : 0);
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_build_grids, dim3(blocks), dim3(256), 0, (hipStream_t)hip_stream, *g, mi_rows, mi_cols,
                       (uint32_t)num_planes, sx, sy, lf_mi, lf_tx_size[0], lf_tx_size[1], lf_tx_size[2]);
    HIP_TRY(hipGetLastError());
    return OXAV1_OK;
}

// OXAV1_OK when all work enqueued on the engine's stream has finished,
// OXAV1_ERR_BUSY while some is still running.  Does not block.
extern "C" int oxav1_pf_engine_query(oxav1_pf_engine *e) {
    if (!e) return OXAV1_ERR_INVALID_ARG;
    const hipError_t q = hipStreamQuery(e->stream);
    if (q == hipSuccess) return OXAV1_OK;
    if (q == hipErrorNotReady) return OXAV1_ERR_BUSY;
    return hip_fail(q, "hipStreamQuery");
}

// --------------------------------------------------------------------- //
// Library basics                                                        //
// --------------------------------------------------------------------- //

extern "C" int oxav1_pf_init(void) { return ensure_device_init(); }

extern "C" const char *oxav1_pf_strerror(int status) {
    switch (status) {
    case OXAV1_OK: return "ok";
    case OXAV1_ERR_UPSCALED_NOT_LARGER: return "superres: upscaled width not larger than frame width";
    case OXAV1_ERR_PLANE_SHAPE_MISMATCH: return "plane shape mismatch";
    case OXAV1_ERR_PLANE_COUNT_MISMATCH: return "plane count mismatch";
    case OXAV1_ERR_INVALID_ARG: return "invalid argument";
    case OXAV1_ERR_ALIAS: return "input and output buffers alias";
    case OXAV1_ERR_HIP: return "HIP runtime error";
    case OXAV1_ERR_OUT_OF_MEMORY: return "out of device memory";
    case OXAV1_ERR_ADMISSION_REFUSED: return "admission refused";
    case OXAV1_ERR_UNSUPPORTED: return "unsupported parameter combination";
    case OXAV1_ERR_NO_DEVICE: return "no HIP device";
    case OXAV1_ERR_BUSY: return "busy: work still in flight";
    case OXAV1_ERR_SCRATCH_TOO_SMALL: return "frame too large for the allocation-free entry point's scratch";
    default: return "unknown status";
    }
}

extern "C" const char *oxav1_pf_last_hip_error(void) { return g_last_hip.c_str(); }

extern "C" uint32_t oxav1_pf_count_units_in_frame(uint32_t unit_size, uint32_t frame_size) {
    if (unit_size == 0) return 1;
    const uint32_t v = (frame_size + (unit_size >> 1)) / unit_size;
    return v > 1 ? v : 1;
}

// Checksum of a constant table of the library (tests compare it with the
// values parsed from the reference): sum of (i+1) * value.
extern "C" int64_t oxav1_pf_table_checksum(const char *name) {
    auto ck = [](const auto &arr) {
        int64_t s = 0;
        size_t i = 0;
        for (auto v : arr) s += (int64_t)(++i) * (int64_t)v;
        return s;
    };
    if (!name) return INT64_MIN;
    std::string n(name);
    if (n == "GAUSSIAN_SEQUENCE") return ck(kGaussianSequence);
    if (n == "UPSCALE_FILTER") return ck(kUpscaleFilter);
    if (n == "SGR_PARAMS") return ck(kSgrParams);
    if (n == "CDEF_DIRECTIONS") return ck(kCdefDirections);
    if (n == "CDEF_UV_DIR") return ck(kCdefUvDir);
    if (n == "NUM_4X4_BLOCKS_WIDE") return ck(kNum4x4W);
    if (n == "NUM_4X4_BLOCKS_HIGH") return ck(kNum4x4H);
    if (n == "TX_WIDTH") return ck(kTxWidth);
    if (n == "TX_HEIGHT") return ck(kTxHeight);
    if (n == "SUBSAMPLED_SIZE") return ck(kSubsampledSize);
    if (n == "MAX_TX_SIZE_RECT") return ck(kMaxTxSizeRect);
    return INT64_MIN;
}

extern "C" int oxav1_pf_dev_alloc(void **ptr, size_t bytes) {
    if (!ptr) return OXAV1_ERR_INVALID_ARG;
    *ptr = nullptr;
    int st = ensure_device_init();
    if (st) return st;
    HIP_TRY(hipMalloc(ptr, bytes ? bytes : 1));
    return OXAV1_OK;
}
extern "C" int oxav1_pf_host_alloc(void **ptr, size_t bytes) {
    if (!ptr) return OXAV1_ERR_INVALID_ARG;
    *ptr = nullptr;
    int st = ensure_device_init();
    if (st) return st;
    HIP_TRY(hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault));
    return OXAV1_OK;
}
extern "C" int oxav1_pf_host_free(void *ptr) {
    if (ptr) HIP_TRY(hipHostFree(ptr));
    return OXAV1_OK;
}
extern "C" int oxav1_pf_dev_free(void *ptr) {
    if (ptr) HIP_TRY(hipFree(ptr));
    return OXAV1_OK;
}
extern "C" int oxav1_pf_dev_upload(void *dst, const void *src, size_t bytes) {
    if (bytes && (!dst || !src)) return OXAV1_ERR_INVALID_ARG;
    if (bytes) HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return OXAV1_OK;
}
extern "C" int oxav1_pf_dev_download(void *dst, const void *src, size_t bytes) {
    if (bytes && (!dst || !src)) return OXAV1_ERR_INVALID_ARG;
    if (bytes) HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return OXAV1_OK;
}
extern "C" int oxav1_pf_dev_fill(void *dst, int byte_value, size_t bytes) {
    if (bytes && !dst) return OXAV1_ERR_INVALID_ARG;
    if (bytes) {
        HIP_TRY(hipMemset(dst, byte_value, bytes));
        HIP_TRY(hipDeviceSynchronize());
    }
    return OXAV1_OK;
}
extern "C" int oxav1_pf_dev_info(int *num_cus, size_t *free_bytes, size_t *total_bytes) {
    int st = ensure_device_init();
    if (st) return st;
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, dev));
    if (num_cus) *num_cus = prop.multiProcessorCount;
    size_t f = 0, t = 0;
    HIP_TRY(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return OXAV1_OK;
}

__global__ void k_copy16(uint4 *__restrict__ dst, const uint4 *__restrict__ src, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

extern "C" int oxav1_pf_dev_copy_kernel(void *dst, const void *src, size_t bytes, int repeats) {
    if (!dst || !src || bytes % 16) return OXAV1_ERR_INVALID_ARG;
    int st = ensure_device_init();
    if (st) return st;
    const size_t n = bytes / 16;
    const uint32_t grid = (uint32_t)std::min<size_t>((n + 255) / 256, 4096);
    for (int r = 0; r < repeats; r++) {
        hipLaunchKernelGGL(k_copy16, dim3(grid ? grid : 1), dim3(256), 0, nullptr, (uint4 *)dst, (const uint4 *)src, n);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipDeviceSynchronize());
    return OXAV1_OK;
}

extern "C" size_t oxav1_pf_struct_size(const char *name) {
    if (!name) return 0;
    std::string n(name);
    if (n == "oxav1_plane_i32") return sizeof(oxav1_plane_i32);
    if (n == "oxav1_lf_ctx") return sizeof(oxav1_lf_ctx);
    if (n == "oxav1_lf_grids") return sizeof(oxav1_lf_grids);
    if (n == "oxav1_cdef_params") return sizeof(oxav1_cdef_params);
    if (n == "oxav1_cdef_ctx") return sizeof(oxav1_cdef_ctx);
    if (n == "oxav1_cdef_grids") return sizeof(oxav1_cdef_grids);
    if (n == "oxav1_sr_ctx") return sizeof(oxav1_sr_ctx);
    if (n == "oxav1_lr_unit") return sizeof(oxav1_lr_unit);
    if (n == "oxav1_lr_ctx") return sizeof(oxav1_lr_ctx);
    if (n == "oxav1_lr_grids") return sizeof(oxav1_lr_grids);
    if (n == "oxav1_film_grain_params") return sizeof(oxav1_film_grain_params);
    if (n == "oxav1_frame_desc") return sizeof(oxav1_frame_desc);
    if (n == "oxav1_frame_dev") return sizeof(oxav1_frame_dev);
    if (n == "oxav1_frame_host") return sizeof(oxav1_frame_host);
    if (n == "oxav1_server_config") return sizeof(oxav1_server_config);
    if (n == "oxav1_tuning") return sizeof(oxav1_tuning);
    if (n == "oxav1_walker_grids") return sizeof(oxav1_walker_grids);
    if (n == "oxav1_server_done") return sizeof(oxav1_server_done);
    if (n == "oxav1_server_stats") return sizeof(oxav1_server_stats);
    if (n == "oxav1_cdef_cand") return sizeof(oxav1_cdef_cand);
    if (n == "oxav1_cdef_ssd_frame") return sizeof(oxav1_cdef_ssd_frame);
    if (n == "oxav1_cdef_elect_input") return sizeof(oxav1_cdef_elect_input);
    if (n == "oxav1_cdef_election") return sizeof(oxav1_cdef_election);
    if (n == "oxav1_lr_elect_input") return sizeof(oxav1_lr_elect_input);
    if (n == "oxav1_lr_unit_cands") return sizeof(oxav1_lr_unit_cands);
    if (n == "oxav1_lr_election") return sizeof(oxav1_lr_election);
    if (n == "oxav1_noise_estimate") return sizeof(oxav1_noise_estimate);
    return 0;
}

#include "dropin.inc"
#include "elect.inc"
#include "lr_elect.inc"
#include "noalloc.inc"
