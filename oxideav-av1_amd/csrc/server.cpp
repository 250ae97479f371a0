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
};

struct Pending {
    uint64_t ticket;
    uint32_t session, slot;
    oxav1_frame_desc desc;
    oxav1_frame_dev dev;
    double arrival;
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
        if (!st) st = oxav1_pf_engine_run(l.eng, 1);
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
