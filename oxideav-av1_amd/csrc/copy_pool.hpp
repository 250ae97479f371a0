// Host-side parallel memcpy (and element-wise conversions) used to stage the
// drop-in calls' pageable planes through pinned buffers.  Host code only (no HIP); shared by engine.hip and
// the CPU fork-safety test (tests/test_copy_pool.py).
#pragma once

#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace oxav1 {

// Each copy is split over up to fifteen worker threads and the calling thread.
// A single-threaded pageable copy runs at one core's memcpy rate; a 4K
// frame's five drop-in calls move about 0.6 GB of int32 samples.
//
// Fork safety: the workers and their mutex and condition variables live in
// one heap object that belongs to the process that started the workers.  A
// forked child never touches it: a condition variable copied by fork still
// counts the parent's waiting workers (so signalling or destroying it waits
// for them forever), and the mutex may have been held at the fork.  The
// child's copies run on the calling thread alone, and its destructor leaves
// the object in place.
class CopyPool {
  public:
    ~CopyPool() {
        Shared *s = s_.load(std::memory_order_acquire);
        if (!s || s->owner != getpid()) return;  // none started, or a forked child
        {
            std::lock_guard<std::mutex> l(s->m);
            s->stop = true;
        }
        s->cv.notify_all();
        for (auto &t : s->th) t.join();
        delete s;
    }

    // dst[0, n) = src[0, n)
    void copy(void *dst, const void *src, size_t n) {
        struct Ctx {
            char *dst;
            const char *src;
        } c{(char *)dst, (const char *)src};
        run(n, 4096, [](void *p, size_t lo, size_t hi) {
            const Ctx &c = *(const Ctx *)p;
            std::memcpy(c.dst + lo, c.src + lo, hi - lo);
        }, &c);
    }

    // fn(ctx, lo, hi) over slices of [0, n) whose bounds are multiples of
    // `align`, on the workers and the calling thread; returns when all are done.
    void run(size_t n, size_t align, void (*fn)(void *, size_t, size_t), void *ctx) {
        const pid_t me = getpid();
        Shared *s = s_.load(std::memory_order_acquire);
        if (s && s->owner != me) {  // forked child
            if (n) fn(ctx, 0, n);
            return;
        }
        if (!s) {
            std::lock_guard<std::mutex> g(start_);
            s = s_.load(std::memory_order_acquire);
            if (!s) {
                s = new Shared;
                s->owner = me;
                const unsigned hc = std::thread::hardware_concurrency();
                const int nw = (int)std::min(15u, hc > 1 ? hc - 1 : 1u);
                for (int i = 0; i < nw; i++) s->th.emplace_back([s, i] { work(s, i + 1); });
                s_.store(s, std::memory_order_release);
            }
        }
        std::unique_lock<std::mutex> l(s->m);
        s->fn = fn;
        s->ctx = ctx;
        s->align = align ? align : 1;
        s->n = n;
        s->parts = (int)s->th.size() + 1;
        s->pending = (int)s->th.size();
        s->gen++;
        l.unlock();
        s->cv.notify_all();
        part(s, 0);
        l.lock();
        s->done.wait(l, [s] { return s->pending == 0; });
    }

  private:
    struct Shared {
        std::mutex m;
        std::condition_variable cv, done;
        std::vector<std::thread> th;
        pid_t owner = 0;
        uint64_t gen = 0;
        int pending = 0, parts = 1;
        bool stop = false;
        void (*fn)(void *, size_t, size_t) = nullptr;
        void *ctx = nullptr;
        size_t align = 1, n = 0;
    };
    static void part(Shared *s, int i) {
        // aligned slices (4 KiB for byte copies) so the threads touch separate pages
        const size_t a = s->align;
        const size_t per = ((s->n + s->parts - 1) / s->parts + a - 1) / a * a;
        const size_t lo = std::min(s->n, per * i), hi = std::min(s->n, lo + per);
        if (hi > lo) s->fn(s->ctx, lo, hi);
    }
    static void work(Shared *s, int i) {
        uint64_t seen = 0;
        std::unique_lock<std::mutex> l(s->m);
        for (;;) {
            s->cv.wait(l, [&] { return s->stop || s->gen != seen; });
            if (s->stop) return;
            seen = s->gen;
            l.unlock();
            part(s, i);
            l.lock();
            if (--s->pending == 0) s->done.notify_one();
        }
    }
    std::atomic<Shared *> s_{nullptr};
    std::mutex start_;  // held only while the first copy starts the workers
};

}  // namespace oxav1
