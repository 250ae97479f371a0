// Serving loop over the chain engine: concurrent decode sessions admitted by
// HBM footprint, one FIFO of ready frames, and lanes (engines on their own
// streams) that take whatever is ready when they free up (continuous
// batching).  Host code only; the filtering itself is the engine's.
//
// The reference decodes one stream per Decoder and filters each frame inline
// (frame_driver.rs:790-1116); sessions, lanes and the arena are this
// library's own serving layer on top of that per-frame contract.

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <vector>

#include "../../include/oxav1_pf.h"

namespace {

constexpr uint64_t kAlign = 256;

uint64_t align_up(uint64_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// First-fit allocator over one device allocation; free ranges coalesce.
struct Arena {
    char *base = nullptr;
    uint64_t size = 0, used = 0;
    std::map<uint64_t, uint64_t> free_;  // offset -> length

    void reset(char *b, uint64_t n) {
        base = b;
        size = n;
        used = 0;
        free_.clear();
        if (n) free_[0] = n;
    }
    bool alloc(uint64_t n, uint64_t *off) {
        n = align_up(n ? n : 1);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            *off = it->first;
            const uint64_t rest = it->second - n, at = it->first + n;
            free_.erase(it);
            if (rest) free_[at] = rest;
            used += n;
            return true;
        }
        return false;
    }
    void release(uint64_t off, uint64_t n) {
        n = align_up(n ? n : 1);
        used -= n;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

uint32_t rnd2(uint32_t x, uint32_t n);

// Where each piece of one frame's inputs sits inside an input buffer
// (oxav1_pf_server_submit_host): the three planes, then the side
// information, 256-byte aligned.
struct InputLayout {
    uint64_t src[3], src_bytes[3], src_rows[3], src_cols[3];
    uint64_t lf_mi, nmi, delta_lf, tx[3], tx_bytes[3], cdef_idx, cdef_bytes, lr[3], lr_bytes[3];
    uint64_t total = 0;
};

// lr_max: room for the most restoration units any frame of this shape can
// signal (every plane restored with 32-sample units), so one layout serves
// all the session's frames.
InputLayout input_layout(const oxav1_frame_desc &d, bool lr_max) {
    InputLayout L{};
    uint64_t at = 0;
    auto take = [&](uint64_t n) {
        const uint64_t o = at;
        at += align_up(n ? n : 1);
        return o;
    };
    const uint64_t B = d.bit_depth == 8 ? 1 : 2;
    const uint32_t sx = d.subsampling_x ? 1 : 0, sy = d.subsampling_y ? 1 : 0;
    for (uint32_t p = 0; p < 3; p++) {
        const uint32_t psx = p ? sx : 0, psy = p ? sy : 0;
        L.src_rows[p] = p < d.num_planes ? (d.mi_rows * 4) >> psy : 0;
        L.src_cols[p] = p < d.num_planes ? (d.mi_cols * 4) >> psx : 0;
        L.src_bytes[p] = L.src_rows[p] * L.src_cols[p] * B;
        L.src[p] = take(L.src_bytes[p]);
    }
    L.nmi = (uint64_t)d.mi_rows * d.mi_cols;
    L.lf_mi = take(L.nmi * 4);
    L.delta_lf = take(L.nmi * 4);
    for (uint32_t p = 0; p < 3; p++) {
        const uint32_t psx = p ? sx : 0, psy = p ? sy : 0;
        L.tx_bytes[p] = (uint64_t)((d.mi_rows + psy) >> psy) * ((d.mi_cols + psx) >> psx);
        L.tx[p] = take(L.tx_bytes[p]);
    }
    L.cdef_bytes = (uint64_t)((d.mi_rows + 15) >> 4) * ((d.mi_cols + 15) >> 4);
    L.cdef_idx = take(L.cdef_bytes);
    for (uint32_t p = 0; p < 3; p++) {
        L.lr_bytes[p] = 0;
        if (p < d.num_planes &&
            (lr_max || (d.lr.uses_lr && d.lr.frame_restoration_type[p] != OXAV1_RESTORE_NONE))) {
            const uint32_t psx = p ? sx : 0, psy = p ? sy : 0;
            const uint32_t us = lr_max ? 32u : d.lr.loop_restoration_size[p];
            const uint32_t ur = oxav1_pf_count_units_in_frame(us, rnd2(d.frame_height, psy));
            const uint32_t uc = oxav1_pf_count_units_in_frame(us, rnd2(std::max(d.upscaled_width, d.frame_width), psx));
            L.lr_bytes[p] = (uint64_t)ur * uc * sizeof(oxav1_lr_unit);
        }
        L.lr[p] = take(L.lr_bytes[p]);
    }
    L.total = at;
    return L;
}

struct Session {
    bool open = false;
    oxav1_frame_desc shape;
    uint32_t nslots = 0;
    uint64_t off = 0, bytes = 0;
    bool grain = false;
    uint32_t rows[3] = {0, 0, 0}, cols[3] = {0, 0, 0};
    std::vector<char *> ref, out;  // nslots * 3 plane pointers
    std::vector<uint8_t> busy;     // per slot: queued or in flight
    uint32_t pending = 0;
    // input region: carved into host-submission input buffers on first use
    char *input = nullptr;
    uint64_t input_bytes = 0;
    std::vector<uint8_t> in_busy;
    std::vector<hipEvent_t> in_ev;  // each buffer's upload
};

struct Pending {
    uint64_t ticket;
    uint32_t session, slot;
    oxav1_frame_desc desc;
    oxav1_frame_dev dev;
    double arrival;
    int inbuf = -1;                   // host submission: the input buffer it occupies
    void *host_ref[3] = {nullptr, nullptr, nullptr}, *host_out[3] = {nullptr, nullptr, nullptr};
};

struct Lane {
    oxav1_pf_engine *eng = nullptr;
    bool busy = false;
    double launch = 0;
    std::vector<Pending> frames;
};

double now_us() {
    using namespace std::chrono;
    return duration<double, std::micro>(steady_clock::now().time_since_epoch()).count();
}

uint32_t rnd2(uint32_t x, uint32_t n) { return n == 0 ? x : (x + (1u << (n - 1))) >> n; }

bool same_shape(const oxav1_frame_desc &a, const oxav1_frame_desc &b) {
    return a.frame_width == b.frame_width && a.frame_height == b.frame_height &&
           a.upscaled_width == b.upscaled_width && a.bit_depth == b.bit_depth &&
           a.subsampling_x == b.subsampling_x && a.subsampling_y == b.subsampling_y &&
           a.num_planes == b.num_planes && a.use_superres == b.use_superres &&
           (a.fg.apply_grain != 0) == (b.fg.apply_grain != 0);
}

}  // namespace

struct oxav1_pf_server {
    int device = 0;
    hipStream_t upload = nullptr;  // host submissions' input copies
    uint32_t max_batch = 16;
    Arena arena;
    void *arena_mem = nullptr;
    uint64_t lane_bytes = 0;
    std::vector<Lane> lanes;
    std::vector<Session> sessions;
    std::deque<Pending> ready;
    std::deque<oxav1_server_done> finished;
    uint64_t next_ticket = 1, submitted = 0, done = 0, launches = 0;
    uint32_t max_batch_seen = 0;
    std::vector<oxav1_frame_desc> dbuf;
    std::vector<oxav1_frame_dev> vbuf;
};

extern "C" double oxav1_pf_server_now_us(void) { return now_us(); }

extern "C" int oxav1_pf_server_destroy(oxav1_pf_server *s) {
    if (!s) return OXAV1_OK;
    for (auto &l : s->lanes)
        if (l.eng) {
            (void)oxav1_pf_engine_sync(l.eng);
            (void)oxav1_pf_engine_destroy(l.eng);
        }
    if (s->upload) (void)hipStreamSynchronize(s->upload);
    for (auto &ss : s->sessions)
        for (auto ev : ss.in_ev)
            if (ev) (void)hipEventDestroy(ev);
    if (s->upload) (void)hipStreamDestroy(s->upload);
    if (s->arena_mem) (void)oxav1_pf_dev_free(s->arena_mem);
    delete s;
    return OXAV1_OK;
}

extern "C" int oxav1_pf_server_create(oxav1_pf_server **out, const oxav1_server_config *cfg) {
    if (!out) return OXAV1_ERR_INVALID_ARG;
    *out = nullptr;
    int st = oxav1_pf_init();
    if (st) return st;
    auto *s = new (std::nothrow) oxav1_pf_server();
    if (!s) return OXAV1_ERR_OUT_OF_MEMORY;
    (void)hipGetDevice(&s->device);
    const uint32_t nl = cfg && cfg->lanes ? cfg->lanes : 4;
    s->max_batch = cfg && cfg->max_batch ? cfg->max_batch : 16;
    if (nl > 16) {
        delete s;
        return OXAV1_ERR_INVALID_ARG;
    }
    s->lanes.resize(nl);
    for (auto &l : s->lanes) {
        st = oxav1_pf_engine_create(&l.eng, nullptr, s->max_batch);
        if (!st && cfg && cfg->largest) {
            uint64_t b = 0;
            st = oxav1_pf_engine_reserve(l.eng, cfg->largest, s->max_batch, &b);
            s->lane_bytes += b;
        }
        if (st) {
            oxav1_pf_server_destroy(s);
            return st;
        }
    }
    uint64_t want = cfg ? cfg->arena_bytes : 0;
    if (!want) {
        int cus = 0;
        size_t fr = 0, tot = 0;
        st = oxav1_pf_dev_info(&cus, &fr, &tot);
        if (st) {
            oxav1_pf_server_destroy(s);
            return st;
        }
        want = fr / 2;
    }
    want = want / kAlign * kAlign;
    st = oxav1_pf_dev_alloc(&s->arena_mem, want);
    if (st) {
        oxav1_pf_server_destroy(s);
        return st;
    }
    s->arena.reset((char *)s->arena_mem, want);
    if (hipStreamCreateWithFlags(&s->upload, hipStreamNonBlocking) != hipSuccess) {
        oxav1_pf_server_destroy(s);
        return OXAV1_ERR_HIP;
    }
    *out = s;
    return OXAV1_OK;
}

extern "C" int oxav1_pf_server_open_session(oxav1_pf_server *s, const oxav1_frame_desc *shape, uint32_t ref_slots,
                                            uint64_t input_bytes, uint32_t *session, void **input,
                                            uint64_t *footprint) {
    if (!s || !shape || !session || ref_slots == 0) return OXAV1_ERR_INVALID_ARG;
    const oxav1_frame_desc &d = *shape;
    if ((d.bit_depth != 8 && d.bit_depth != 10 && d.bit_depth != 12) || (d.num_planes != 1 && d.num_planes != 3) ||
        d.frame_height == 0 || d.frame_width == 0)
        return OXAV1_ERR_INVALID_ARG;
    Session ss;
    ss.open = true;
    ss.shape = d;
    ss.nslots = ref_slots;
    ss.grain = d.fg.apply_grain != 0;
    const uint32_t uw = std::max(d.upscaled_width, d.frame_width);
    const uint64_t B = d.bit_depth == 8 ? 1 : 2;
    uint64_t plane_bytes[3] = {0, 0, 0}, slot_bytes = 0;
    for (uint32_t p = 0; p < d.num_planes; p++) {
        const uint32_t psx = p && d.subsampling_x ? 1 : 0, psy = p && d.subsampling_y ? 1 : 0;
        ss.rows[p] = rnd2(d.frame_height, psy);
        ss.cols[p] = rnd2(uw, psx);
        plane_bytes[p] = align_up((uint64_t)ss.rows[p] * ss.cols[p] * B);
        slot_bytes += plane_bytes[p];
    }
    if (ss.grain) slot_bytes *= 2;
    ss.bytes = slot_bytes * ref_slots + align_up(input_bytes);
    if (!s->arena.alloc(ss.bytes, &ss.off)) return OXAV1_ERR_ADMISSION_REFUSED;
    char *at = s->arena.base + ss.off;
    ss.ref.assign((size_t)ref_slots * 3, nullptr);
    ss.out.assign((size_t)ref_slots * 3, nullptr);
    ss.busy.assign(ref_slots, 0);
    for (uint32_t k = 0; k < ref_slots; k++)
        for (int w = 0; w < (ss.grain ? 2 : 1); w++)
            for (uint32_t p = 0; p < d.num_planes; p++) {
                (w ? ss.out : ss.ref)[k * 3 + p] = at;
                at += plane_bytes[p];
            }
    if (input) *input = input_bytes ? at : nullptr;
    ss.input = input_bytes ? at : nullptr;
    ss.input_bytes = input_bytes;
    if (footprint) *footprint = ss.bytes;
    uint32_t id = 0;
    while (id < s->sessions.size() && s->sessions[id].open) id++;
    if (id == s->sessions.size()) s->sessions.emplace_back();
    s->sessions[id] = std::move(ss);
    *session = id;
    return OXAV1_OK;
}

extern "C" int oxav1_pf_server_close_session(oxav1_pf_server *s, uint32_t session) {
    if (!s || session >= s->sessions.size() || !s->sessions[session].open) return OXAV1_ERR_INVALID_ARG;
    Session &ss = s->sessions[session];
    if (ss.pending) return OXAV1_ERR_BUSY;
    s->arena.release(ss.off, ss.bytes);
    for (auto ev : ss.in_ev)
        if (ev) (void)hipEventDestroy(ev);
    ss = Session();
    return OXAV1_OK;
}

extern "C" int oxav1_pf_server_slot(oxav1_pf_server *s, uint32_t session, uint32_t slot, uint32_t plane, int which,
                                    void **ptr, uint32_t *pitch) {
    if (!s || !ptr || session >= s->sessions.size() || !s->sessions[session].open) return OXAV1_ERR_INVALID_ARG;
    const Session &ss = s->sessions[session];
    if (slot >= ss.nslots || plane >= ss.shape.num_planes || (which && !ss.grain)) return OXAV1_ERR_INVALID_ARG;
    *ptr = (which ? ss.out : ss.ref)[slot * 3 + plane];
    if (pitch) *pitch = ss.cols[plane];
    return OXAV1_OK;
}

extern "C" int oxav1_pf_server_submit(oxav1_pf_server *s, uint32_t session, const oxav1_frame_desc *desc,
                                      const oxav1_frame_dev *dev, uint32_t slot, double arrival_us,
                                      uint64_t *ticket) {
    if (!s || !desc || !dev || session >= s->sessions.size() || !s->sessions[session].open)
        return OXAV1_ERR_INVALID_ARG;
    Session &ss = s->sessions[session];
    if (slot >= ss.nslots || !same_shape(*desc, ss.shape)) return OXAV1_ERR_INVALID_ARG;
    if (ss.busy[slot]) return OXAV1_ERR_BUSY;
    Pending f;
    f.ticket = s->next_ticket++;
    f.session = session;
    f.slot = slot;
    f.desc = *desc;
    f.dev = *dev;
    for (uint32_t p = 0; p < 3; p++) {
        const bool on = p < ss.shape.num_planes;
        f.dev.ref_out[p] = on ? ss.ref[slot * 3 + p] : nullptr;
        f.dev.ref_pitch[p] = on ? ss.cols[p] : 0;
        f.dev.out[p] = on && ss.grain ? ss.out[slot * 3 + p] : nullptr;
        f.dev.out_pitch[p] = on && ss.grain ? ss.cols[p] : 0;
    }
    f.arrival = arrival_us;
    ss.busy[slot] = 1;
    ss.pending++;
    s->ready.push_back(f);
    s->submitted++;
    if (ticket) *ticket = f.ticket;
    return OXAV1_OK;
}

extern "C" int oxav1_pf_server_submit_existing(oxav1_pf_server *s, uint32_t session, uint32_t slot,
                                               const oxav1_film_grain_params *fg, double arrival_us,
                                               uint64_t *ticket) {
    if (!s || !fg || session >= s->sessions.size() || !s->sessions[session].open) return OXAV1_ERR_INVALID_ARG;
    Session &ss = s->sessions[session];
    if (slot >= ss.nslots || !ss.grain) return OXAV1_ERR_INVALID_ARG;
    if (ss.busy[slot]) return OXAV1_ERR_BUSY;
    // Nothing before the grain stage runs; the stored frame is both the
    // source and the pre-grain reference, the slot's grain buffer the output.
    oxav1_frame_desc d = ss.shape;
    d.apply_lf = 0;
    d.apply_cdef = 0;
    d.use_superres = 0;
    d.frame_width = std::max(d.upscaled_width, d.frame_width);
    d.upscaled_width = d.frame_width;
    d.lr.uses_lr = 0;
    d.fg = *fg;
    d.mi_rows = (d.frame_height + 3) / 4;  // the mi grid of the stored (upscaled) frame
    d.mi_cols = (d.frame_width + 3) / 4;
    Pending f;
    f.ticket = s->next_ticket++;
    f.session = session;
    f.slot = slot;
    f.desc = d;
    std::memset(&f.dev, 0, sizeof f.dev);
    for (uint32_t p = 0; p < ss.shape.num_planes; p++) {
        f.dev.src[p] = ss.ref[slot * 3 + p];
        f.dev.src_pitch[p] = ss.cols[p];
        f.dev.ref_out[p] = ss.ref[slot * 3 + p];
        f.dev.ref_pitch[p] = ss.cols[p];
        f.dev.out[p] = ss.out[slot * 3 + p];
        f.dev.out_pitch[p] = ss.cols[p];
    }
    f.arrival = arrival_us;
    ss.busy[slot] = 1;
    ss.pending++;
    s->ready.push_back(f);
    s->submitted++;
    if (ticket) *ticket = f.ticket;
    return OXAV1_OK;
}

extern "C" uint64_t oxav1_pf_server_frame_input_bytes(const oxav1_frame_desc *shape) {
    return shape ? input_layout(*shape, true).total : 0;
}

extern "C" int oxav1_pf_server_submit_host(oxav1_pf_server *s, uint32_t session, const oxav1_frame_desc *desc,
                                           const oxav1_frame_host *in, uint32_t slot, double arrival_us,
                                           uint64_t *ticket) {
    if (!s || !desc || !in || session >= s->sessions.size() || !s->sessions[session].open)
        return OXAV1_ERR_INVALID_ARG;
    Session &ss = s->sessions[session];
    if (slot >= ss.nslots || !same_shape(*desc, ss.shape)) return OXAV1_ERR_INVALID_ARG;
    // the session's fixed layout (offsets, stride) and this frame's sizes
    const InputLayout L = input_layout(ss.shape, true), F = input_layout(*desc, false);
    if (!ss.input || L.total > ss.input_bytes || F.nmi != L.nmi) return OXAV1_ERR_INVALID_ARG;
    for (uint32_t p = 0; p < 3; p++)
        if (F.lr_bytes[p] > L.lr_bytes[p] || F.src_bytes[p] != L.src_bytes[p]) return OXAV1_ERR_INVALID_ARG;
    if (!in->lf_mi || !in->cdef_idx || !in->lf_tx_size[0]) return OXAV1_ERR_INVALID_ARG;
    for (uint32_t p = 0; p < desc->num_planes; p++)
        if (!in->src[p] || (p && !in->lf_tx_size[p]) || (F.lr_bytes[p] && !in->lr_units[p]))
            return OXAV1_ERR_INVALID_ARG;
    if (ss.busy[slot]) return OXAV1_ERR_BUSY;
    if (ss.in_busy.empty()) {  // carve the input region on first use
        const uint64_t n = ss.input_bytes / L.total;
        ss.in_busy.assign(n, 0);
        ss.in_ev.assign(n, nullptr);
        for (auto &ev : ss.in_ev)
            if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) return OXAV1_ERR_HIP;
    }
    int b = -1;
    for (size_t i = 0; i < ss.in_busy.size(); i++)
        if (!ss.in_busy[i]) {
            b = (int)i;
            break;
        }
    if (b < 0) return OXAV1_ERR_BUSY;
    char *base = ss.input + (uint64_t)b * L.total;
    const uint64_t B = desc->bit_depth == 8 ? 1 : 2;
    hipStream_t u = s->upload;
    Pending f;
    std::memset(&f.dev, 0, sizeof f.dev);
    for (uint32_t p = 0; p < desc->num_planes; p++) {
        char *dst = base + L.src[p];
        const uint64_t row = L.src_cols[p] * B, pitch = in->src_pitch[p] ? in->src_pitch[p] * B : row;
        if (pitch == row) {
            if (hipMemcpyAsync(dst, in->src[p], L.src_bytes[p], hipMemcpyHostToDevice, u) != hipSuccess)
                return OXAV1_ERR_HIP;
        } else if (hipMemcpy2DAsync(dst, row, in->src[p], pitch, row, L.src_rows[p], hipMemcpyHostToDevice, u) !=
                   hipSuccess) {
            return OXAV1_ERR_HIP;
        }
        f.dev.src[p] = dst;
        f.dev.src_pitch[p] = (uint32_t)L.src_cols[p];
    }
    auto put = [&](uint64_t off, const void *src, uint64_t n) -> bool {
        return !n || hipMemcpyAsync(base + off, src, n, hipMemcpyHostToDevice, u) == hipSuccess;
    };
    bool ok = put(L.lf_mi, in->lf_mi, L.nmi * 4) && put(L.cdef_idx, in->cdef_idx, L.cdef_bytes);
    if (in->delta_lf) ok = ok && put(L.delta_lf, in->delta_lf, L.nmi * 4);
    for (uint32_t p = 0; p < 3 && ok; p++) {
        const bool on = p < desc->num_planes;
        if (on) ok = put(L.tx[p], in->lf_tx_size[p], L.tx_bytes[p]);
        if (F.lr_bytes[p]) ok = ok && put(L.lr[p], in->lr_units[p], F.lr_bytes[p]);
    }
    if (!ok || hipEventRecord(ss.in_ev[b], u) != hipSuccess) return OXAV1_ERR_HIP;
    f.dev.lf_mi = (const uint32_t *)(base + L.lf_mi);
    f.dev.delta_lf = in->delta_lf ? (const int8_t *)(base + L.delta_lf) : nullptr;
    for (uint32_t p = 0; p < 3; p++) {
        f.dev.lf_tx_size[p] = p < desc->num_planes ? (const uint8_t *)(base + L.tx[p]) : nullptr;
        f.dev.lr_units[p] = F.lr_bytes[p] ? (const oxav1_lr_unit *)(base + L.lr[p]) : nullptr;
    }
    f.dev.cdef_idx = (const int8_t *)(base + L.cdef_idx);
    f.ticket = s->next_ticket++;
    f.session = session;
    f.slot = slot;
    f.desc = *desc;
    for (uint32_t p = 0; p < 3; p++) {
        const bool on = p < ss.shape.num_planes;
        f.dev.ref_out[p] = on ? ss.ref[slot * 3 + p] : nullptr;
        f.dev.ref_pitch[p] = on ? ss.cols[p] : 0;
        f.dev.out[p] = on && ss.grain ? ss.out[slot * 3 + p] : nullptr;
        f.dev.out_pitch[p] = on && ss.grain ? ss.cols[p] : 0;
        f.host_ref[p] = on ? in->ref_out[p] : nullptr;
        f.host_out[p] = on && ss.grain ? in->out[p] : nullptr;
    }
    f.arrival = arrival_us;
    f.inbuf = b;
    ss.in_busy[b] = 1;
    ss.busy[slot] = 1;
    ss.pending++;
    s->ready.push_back(f);
    s->submitted++;
    if (ticket) *ticket = f.ticket;
    return OXAV1_OK;
}

static void retire(oxav1_pf_server *s, Lane &l, int status, double t) {
    const uint32_t n = (uint32_t)l.frames.size();
    for (const Pending &f : l.frames) {
        oxav1_server_done r;
        r.ticket = f.ticket;
        r.session = f.session;
        r.slot = f.slot;
        r.status = status;
        r.batch = n;
        r.arrival_us = f.arrival;
        r.launch_us = l.launch;
        r.done_us = t;
        s->finished.push_back(r);
        Session &ss = s->sessions[f.session];
        ss.busy[f.slot] = 0;
        if (f.inbuf >= 0) ss.in_busy[f.inbuf] = 0;
        ss.pending--;
        s->done++;
    }
    l.frames.clear();
    l.busy = false;
}

extern "C" int oxav1_pf_server_poll(oxav1_pf_server *s, oxav1_server_done *done, uint32_t cap, uint32_t *n_done) {
    if (!s || (cap && !done)) return OXAV1_ERR_INVALID_ARG;
    int first_err = OXAV1_OK;
    // 1. retire lanes whose stream has drained
    for (auto &l : s->lanes) {
        if (!l.busy) continue;
        const int q = oxav1_pf_engine_query(l.eng);
        if (q == OXAV1_ERR_BUSY) continue;
        retire(s, l, q, now_us());
        if (q && !first_err) first_err = q;
    }
    // 2. hand what is ready now to free lanes, one batch per lane
    for (auto &l : s->lanes) {
        if (s->ready.empty()) break;
        if (l.busy) continue;
        const uint32_t n = (uint32_t)std::min<size_t>(s->ready.size(), s->max_batch);
        s->dbuf.resize(n);
        s->vbuf.resize(n);
        l.frames.clear();
        for (uint32_t i = 0; i < n; i++) {
            l.frames.push_back(s->ready.front());
            s->ready.pop_front();
            s->dbuf[i] = l.frames.back().desc;
            s->vbuf[i] = l.frames.back().dev;
        }
        l.launch = now_us();
        int st = oxav1_pf_engine_prepare(l.eng, s->dbuf.data(), s->vbuf.data(), n);
        hipStream_t es = (hipStream_t)oxav1_pf_engine_stream(l.eng);
        // host submissions: the chain waits for the frames' input copies
        for (const Pending &f : l.frames)
            if (!st && f.inbuf >= 0 &&
                hipStreamWaitEvent(es, s->sessions[f.session].in_ev[f.inbuf], 0) != hipSuccess)
                st = OXAV1_ERR_HIP;
        if (!st) st = oxav1_pf_engine_run(l.eng, 1);
        // results copied back to the host destinations before the lane drains
        for (const Pending &f : l.frames) {
            const Session &ss = s->sessions[f.session];
            const uint64_t B = ss.shape.bit_depth == 8 ? 1 : 2;
            for (uint32_t p = 0; p < ss.shape.num_planes && !st; p++) {
                const uint64_t n = (uint64_t)ss.rows[p] * ss.cols[p] * B;
                if (f.host_ref[p] &&
                    hipMemcpyAsync(f.host_ref[p], f.dev.ref_out[p], n, hipMemcpyDeviceToHost, es) != hipSuccess)
                    st = OXAV1_ERR_HIP;
                if (f.host_out[p] && f.dev.out[p] &&
                    hipMemcpyAsync(f.host_out[p], f.dev.out[p], n, hipMemcpyDeviceToHost, es) != hipSuccess)
                    st = OXAV1_ERR_HIP;
            }
        }
        if (st) {
            retire(s, l, st, now_us());
            if (!first_err) first_err = st;
            continue;
        }
        l.busy = true;
        s->launches++;
        s->max_batch_seen = std::max(s->max_batch_seen, n);
    }
    // 3. hand back finished frames
    uint32_t k = 0;
    while (k < cap && !s->finished.empty()) {
        done[k++] = s->finished.front();
        s->finished.pop_front();
    }
    if (n_done) *n_done = k;
    return first_err;
}

extern "C" int oxav1_pf_server_stats(oxav1_pf_server *s, oxav1_server_stats *st) {
    if (!s || !st) return OXAV1_ERR_INVALID_ARG;
    std::memset(st, 0, sizeof *st);
    for (const auto &ss : s->sessions) st->sessions_open += ss.open ? 1 : 0;
    for (const auto &l : s->lanes) st->lanes_busy += l.busy ? 1 : 0;
    st->frames_queued = (uint32_t)s->ready.size();
    st->max_batch_seen = s->max_batch_seen;
    st->frames_submitted = s->submitted;
    st->frames_done = s->done;
    st->launches = s->launches;
    st->arena_bytes = s->arena.size;
    st->arena_used = s->arena.used;
    st->lane_bytes = s->lane_bytes;
    return OXAV1_OK;
}
