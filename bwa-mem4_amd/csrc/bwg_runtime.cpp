// Host runtime: device memory pool, stream sets, batch staging, the plain form's launches, result
// copies.  The batched form's planning and launches are in bwg_batched.cpp.
#include "bwg_runtime.h"

#include <sched.h>

#include <algorithm>
#include <cstring>
#include <deque>
#include <iterator>
#include <new>
#include <system_error>

namespace bwg {

// ---------------------------------------------------------------------------------------------
// DevicePool
// ---------------------------------------------------------------------------------------------
size_t DevicePool::round(size_t b) {
    if (b < 4096) return 4096;
    size_t r = 4096;
    while (r < b) r <<= 1;
    if (r - b > (r >> 2) && b > (1u << 20)) r = (b + (1u << 20) - 1) & ~(size_t)((1u << 20) - 1);
    return r;
}

void *DevicePool::alloc(size_t bytes) {
    const size_t sz = round(bytes);
    std::lock_guard<std::mutex> g(mu_);
    auto it = free_.lower_bound(sz);
    void *p = nullptr;
    size_t got = sz;
    if (it != free_.end() && it->first <= sz + (sz >> 2)) {
        p = it->second;
        got = it->first;
        free_.erase(it);
        cached_ -= got;
    } else if (hipMalloc(&p, sz) != hipSuccess) {
        // drop the cache and retry once
        (void)hipGetLastError();
        trim_locked(0);
        if (hipMalloc(&p, sz) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
    }
    live_[p] = got;
    in_use_ += got;
    if (in_use_.load() > peak_.load()) peak_.store(in_use_.load());
    return p;
}

void DevicePool::free(void *p) {
    if (!p) return;
    std::lock_guard<std::mutex> g(mu_);
    auto it = live_.find(p);
    if (it == live_.end()) return;
    free_.emplace(it->second, p);
    cached_ += it->second;
    in_use_ -= it->second;
    live_.erase(it);
    if (cached_.load() > cap_.load()) trim_locked(cap_.load());
}

void DevicePool::trim_locked(size_t cap) {
    // largest blocks first: they return the most memory per hipFree
    while (!free_.empty() && cached_.load() > cap) {
        auto it = std::prev(free_.end());
        (void)hipFree(it->second);
        cached_ -= it->first;
        free_.erase(it);
    }
}

void DevicePool::set_cache_cap(size_t cap) {
    std::lock_guard<std::mutex> g(mu_);
    cap_ = cap;
    trim_locked(cap);
}

void DevicePool::release_all() {
    std::lock_guard<std::mutex> g(mu_);
    for (auto &kv : free_) (void)hipFree(kv.second);
    for (auto &kv : live_) (void)hipFree(kv.first);
    free_.clear();
    live_.clear();
    in_use_ = 0;
    cached_ = 0;
}

// ---------------------------------------------------------------------------------------------
// StreamSet
// ---------------------------------------------------------------------------------------------
namespace {

// The host worker threads behind run_parts: started on first use, kept for the life of the
// process (a thread start costs tens of microseconds, and a large call runs several passes).
// Calls from several threads share the workers; each call's parts are claimed with an atomic
// counter, by the workers and by the calling thread alike.  A worker counts itself into a call
// (users) while it may touch it, and the caller returns only after taking the call out of the
// queue with no worker inside it.
class HostWorkers {
  public:
    static HostWorkers &get() {
        static HostWorkers *w = new HostWorkers();  // never destroyed: workers run until exit
        return *w;
    }
    void run(int parts, const std::function<void(int)> &fn) {
        Call call;
        call.fn = &fn;
        call.parts = parts;
        call.left.store(parts);
        {
            std::lock_guard<std::mutex> g(mu_);
            calls_.push_back(&call);
        }
        for (int q = 1; q < parts; ++q) cv_.notify_one();  // as many workers as the call can use
        work_on(call);
        std::unique_lock<std::mutex> g(mu_);
        const auto it = std::find(calls_.begin(), calls_.end(), &call);
        if (it != calls_.end()) calls_.erase(it);
        done_.wait(g, [&] { return call.users == 0 && call.left.load() == 0; });
    }

  private:
    struct Call {
        const std::function<void(int)> *fn = nullptr;
        int parts = 0;
        int users = 0;  // workers inside the call (under mu_)
        std::atomic<int> next{0}, left{0};
    };
    HostWorkers() {
        for (int k = 0; k < 15; ++k) {
            try {
                std::thread([this] { loop(); }).detach();
            } catch (const std::system_error &) {
                break;  // fewer workers: the callers run more of their parts themselves
            }
        }
    }
    // claim and run parts of `call` until none is left to claim
    static void work_on(Call &call) {
        for (;;) {
            const int q = call.next.fetch_add(1);
            if (q >= call.parts) return;
            (*call.fn)(q);
            call.left.fetch_sub(1);
        }
    }
    void loop() {
        for (;;) {
            Call *call = nullptr;
            {
                std::unique_lock<std::mutex> g(mu_);
                cv_.wait(g, [&] { return !calls_.empty(); });
                call = calls_.front();
                ++call->users;
                // every part claimed: no one else needs to find the call
                if (call->next.load() >= call->parts) calls_.pop_front();
            }
            work_on(*call);
            {
                std::lock_guard<std::mutex> g(mu_);
                --call->users;
            }
            done_.notify_all();
        }
    }
    std::mutex mu_;
    std::condition_variable cv_, done_;
    std::deque<Call *> calls_;
};

}  // namespace

void run_parts(int parts, const std::function<void(int)> &fn) {
    if (parts <= 1) {
        if (parts == 1) fn(0);
        return;
    }
    HostWorkers::get().run(parts, fn);
}

int host_parts(const Tuning &tu, size_t n) {
    static const int avail = [] {
        cpu_set_t set;
        CPU_ZERO(&set);
        int k = sched_getaffinity(0, sizeof set, &set) == 0 ? CPU_COUNT(&set) : 0;
        if (k <= 0) k = (int)std::thread::hardware_concurrency();
        return std::max(1, k);
    }();
    // an explicit thread count also splits smaller calls (down to 1024 jobs a thread)
    const int cap = tu.host_threads > 0 ? tu.host_threads : std::min(avail, 16);
    const size_t grain = tu.host_threads > 0 ? 1024 : kHostGrain;
    return (int)std::max<size_t>(1, std::min<size_t>((size_t)cap, n / grain));
}

int StreamSet::create(int rotation) {
    hipStream_t s[kAuxStreams + 1] = {};
    for (int k = 0; k <= kAuxStreams; ++k)
        if (hipStreamCreateWithFlags(&s[k], hipStreamNonBlocking) != hipSuccess) {
            for (int r = 0; r < k; ++r) (void)hipStreamDestroy(s[r]);
            return BWG_EIO;
        }
    const int m = ((rotation % (kAuxStreams + 1)) + kAuxStreams + 1) % (kAuxStreams + 1);
    main = s[m];
    for (int k = 0, a = 0; k <= kAuxStreams; ++k)
        if (k != m) aux[a++] = s[k];
    if (hipEventCreateWithFlags(&fork, hipEventDisableTiming) != hipSuccess) return BWG_EIO;
    for (int k = 0; k < kAuxStreams; ++k)
        if (hipEventCreateWithFlags(&join[k], hipEventDisableTiming) != hipSuccess) return BWG_EIO;
    return BWG_OK;
}

void StreamSet::destroy() {
    if (main) (void)hipStreamSynchronize(main);
    for (int k = 0; k < kAuxStreams; ++k) {
        if (aux[k]) {
            (void)hipStreamSynchronize(aux[k]);
            (void)hipStreamDestroy(aux[k]);
        }
        if (join[k]) (void)hipEventDestroy(join[k]);
        aux[k] = nullptr;
        join[k] = nullptr;
    }
    if (fork) (void)hipEventDestroy(fork);
    if (main) (void)hipStreamDestroy(main);
    main = nullptr;
    fork = nullptr;
}

int ExecLane::create(int index) { return ss.create(index + 1); }  // the context's own set took rotation 0

void ExecLane::destroy() {
    ss.destroy();
    if (pin_out) (void)hipHostFree(pin_out);
    pin_out = nullptr;
    pin_out_cap = 0;
}

ExecLane *acquire_lane(bwg_ctx *c) {
    std::unique_lock<std::mutex> g(c->lane_mu);
    for (;;) {
        if (!c->free_lanes.empty()) {
            ExecLane *l = c->free_lanes.back();
            c->free_lanes.pop_back();
            return l;
        }
        if ((int)c->lanes.size() < c->max_lanes) {
            std::unique_ptr<ExecLane> l(new (std::nothrow) ExecLane());
            if (!l) return nullptr;
            if (l->create((int)c->lanes.size()) != BWG_OK) {
                l->destroy();
                return nullptr;
            }
            c->lanes.push_back(std::move(l));
            return c->lanes.back().get();
        }
        c->lane_cv.wait(g);
    }
}

void release_lane(bwg_ctx *c, ExecLane *l) {
    {
        std::lock_guard<std::mutex> g(c->lane_mu);
        c->free_lanes.push_back(l);
    }
    c->lane_cv.notify_one();
}

// ---------------------------------------------------------------------------------------------
// staging
// ---------------------------------------------------------------------------------------------
namespace {

// A batch object; with `ev` its timing events are borrowed (the allocation-free path), else its own.
bwg_batch *new_batch(const bwg_ctx *c, int kind, size_t n, const hipEvent_t *ev = nullptr) {
    bwg_batch *b = new (std::nothrow) bwg_batch();
    if (!b) return nullptr;
    b->kind = kind;
    b->n = n;
    b->tu = c->tuning_snapshot();
    if (ev) {
        b->ev_start = ev[0];
        b->ev_end = ev[1];
        b->borrowed_events = true;
        return b;
    }
    if (hipEventCreate(&b->ev_start) != hipSuccess || hipEventCreate(&b->ev_end) != hipSuccess) {
        if (b->ev_start) (void)hipEventDestroy(b->ev_start);
        delete b;
        return nullptr;
    }
    return b;
}

// Arenas from this size on are uploaded from a helper thread when the caller defers them.
constexpr size_t kDeferArenaBytes = 8u << 20;

int stage_buffers(bwg_ctx *c, bwg_batch *b, const uint8_t *arena, size_t arena_len, const void *jobs,
                  size_t job_bytes, size_t out_bytes, hipStream_t st, bool defer_arena = false) {
    b->arena_len = arena_len;
    b->d_arena = (uint8_t *)batch_alloc(c, b, arena_len + 64);
    b->d_jobs = batch_alloc(c, b, std::max<size_t>(job_bytes, 64));
    b->d_out = batch_alloc(c, b, std::max<size_t>(out_bytes, 64));
    if (!b->d_arena || !b->d_jobs || !b->d_out) return BWG_ENOMEM;
    hipError_t e = hipSuccess;
    if (job_bytes) e = hipMemcpyAsync(b->d_jobs, jobs, job_bytes, hipMemcpyHostToDevice, st);
    if (e != hipSuccess || !arena_len) return bwg_map_hip(e);
    if (defer_arena && arena_len >= kDeferArenaBytes && !b->bump) {
        // an upload from pageable memory keeps its host threads busy for the whole transfer: give
        // it threads of its own, in stream order on st, while the caller plans; it goes through
        // the context's pinned chunks (several copying threads) when a set is free
        const int dev = c->device;
        uint8_t *dst = b->d_arena;
        try {
            b->arena_copy = std::thread([c, b, dev, dst, arena, arena_len, st] {
                hipError_t r = hipSetDevice(dev);
                UploadStaging *u = r == hipSuccess ? acquire_upload(c) : nullptr;
                if (u) {
                    r = upload_pinned(*u, dst, arena, arena_len, st, dev);
                    release_upload(c, u);
                } else if (r == hipSuccess) {
                    r = hipMemcpyAsync(dst, arena, arena_len, hipMemcpyHostToDevice, st);
                }
                b->arena_err = r;
            });
            return BWG_OK;
        } catch (const std::system_error &) {
            // no thread to spare: upload here
        }
    }
    return bwg_map_hip(hipMemcpyAsync(b->d_arena, arena, arena_len, hipMemcpyHostToDevice, st));
}

}  // namespace

void *batch_alloc(bwg_ctx *c, bwg_batch *b, size_t bytes) {
    return b->bump ? b->bump->take(bytes) : c->pool.alloc(bytes);
}

int wait_arena(bwg_batch *b) {
    if (!b->arena_copy.joinable()) return BWG_OK;
    b->arena_copy.join();
    return bwg_map_hip(b->arena_err);
}

int UploadStaging::create() {
    for (int k = 0; k < kBufs; ++k) {
        if (hipHostMalloc(&buf[k], kChunk, hipHostMallocDefault) != hipSuccess) {
            buf[k] = nullptr;
            return BWG_ENOMEM;
        }
        if (hipEventCreateWithFlags(&ev[k], hipEventDisableTiming) != hipSuccess) {
            ev[k] = nullptr;
            return BWG_EIO;
        }
    }
    return BWG_OK;
}

void UploadStaging::destroy() {
    for (int k = 0; k < kBufs; ++k) {
        if (ev[k]) {
            if (pending[k]) (void)hipEventSynchronize(ev[k]);
            (void)hipEventDestroy(ev[k]);
        }
        if (buf[k]) (void)hipHostFree(buf[k]);
        ev[k] = nullptr;
        buf[k] = nullptr;
        pending[k] = false;
    }
}

UploadStaging *acquire_upload(bwg_ctx *c) {
    std::lock_guard<std::mutex> g(c->up_mu);
    if (!c->up_free.empty()) {
        UploadStaging *u = c->up_free.back();
        c->up_free.pop_back();
        return u;
    }
    if ((int)c->up_all.size() >= bwg_ctx::kMaxUploadSets) return nullptr;
    std::unique_ptr<UploadStaging> u(new (std::nothrow) UploadStaging());
    if (!u) return nullptr;
    if (u->create() != BWG_OK) {
        u->destroy();
        return nullptr;
    }
    c->up_all.push_back(std::move(u));
    return c->up_all.back().get();
}

void release_upload(bwg_ctx *c, UploadStaging *u) {
    std::lock_guard<std::mutex> g(c->up_mu);
    c->up_free.push_back(u);
}

void destroy_uploads(bwg_ctx *c) {
    std::lock_guard<std::mutex> g(c->up_mu);
    for (auto &u : c->up_all) u->destroy();
    c->up_all.clear();
    c->up_free.clear();
}

hipError_t upload_pinned(UploadStaging &u, uint8_t *dst, const uint8_t *src, size_t len, hipStream_t st,
                         int device) {
    const size_t chunks = (len + UploadStaging::kChunk - 1) / UploadStaging::kChunk;
    const int T = (int)std::max<size_t>(1, std::min<size_t>(UploadStaging::kThreads, chunks));
    std::atomic<int> err{(int)hipSuccess};
    // thread t copies chunks t, t + T, ... through its own buffers 2t and 2t + 1 in turn
    auto worker = [&](int t) {
        hipError_t e = hipSetDevice(device);
        int flip = 0;
        for (size_t ch = (size_t)t; ch < chunks && e == hipSuccess && err.load() == (int)hipSuccess; ch += (size_t)T) {
            const int bi = 2 * t + flip;
            flip ^= 1;
            if (u.pending[bi]) e = hipEventSynchronize(u.ev[bi]);  // its last DMA has read it
            if (e != hipSuccess) break;
            const size_t off = ch * UploadStaging::kChunk;
            const size_t sz = std::min(UploadStaging::kChunk, len - off);
            std::memcpy(u.buf[bi], src + off, sz);
            e = hipMemcpyAsync(dst + off, u.buf[bi], sz, hipMemcpyHostToDevice, st);
            if (e == hipSuccess) e = hipEventRecord(u.ev[bi], st);
            if (e == hipSuccess) u.pending[bi] = true;
        }
        if (e != hipSuccess) err.store((int)e);
    };
    std::vector<std::thread> th;
    try {
        for (int t = 1; t < T; ++t) th.emplace_back(worker, t);
    } catch (const std::system_error &) {
        err.store((int)hipErrorOutOfMemory);
    }
    worker(0);
    for (auto &t : th) t.join();
    return (hipError_t)err.load();
}

namespace {

int stage_ext_impl(bwg_ctx *c, const ExtScoring &sc, size_t n, const uint8_t *arena, size_t arena_len,
                   const uint64_t *q_off, const uint32_t *q_len, const uint64_t *t_off, const uint32_t *t_len,
                   const int32_t *h0, const uint8_t *ref, Bump *bump, const hipEvent_t *ev, hipStream_t st,
                   bwg_batch **out, bool defer_arena) {
    if (hipSetDevice(c->device) != hipSuccess) return BWG_EIO;
    bwg_batch *b = new_batch(c, 0, n, ev);
    if (!b) return BWG_ENOMEM;
    b->bump = bump;
    b->esc = sc;
    b->d_ref = ref;
    b->hext.resize(n);
    parallel_for(n, host_parts(b->tu, n), [&](size_t lo, size_t hi, int) {
        for (size_t k = lo; k < hi; ++k) {
            ExtJob &j = b->hext[k];
            j.q_off = q_off[k];
            j.t_off = t_off[k];
            j.qlen = q_len[k];
            j.tlen = t_len[k];
            j.h0 = h0[k];
            j.aux = (uint32_t)k;
        }
    });
    int rc = stage_buffers(c, b, arena, arena_len, b->hext.data(), n * sizeof(ExtJob),
                           n * sizeof(bwg_ext_result), st, defer_arena);
    if (rc) {
        destroy_batch(c, b);
        return rc;
    }
    *out = b;
    return BWG_OK;
}

}  // namespace

int stage_ext(bwg_ctx *c, const ExtScoring &sc, size_t n, const uint8_t *arena, size_t arena_len,
              const uint64_t *q_off, const uint32_t *q_len, const uint64_t *t_off,
              const uint32_t *t_len, const int32_t *h0, const uint8_t *ref, hipStream_t st,
              bwg_batch **out, bool defer_arena) {
    return stage_ext_impl(c, sc, n, arena, arena_len, q_off, q_len, t_off, t_len, h0, ref, nullptr, nullptr, st,
                          out, defer_arena);
}

int stage_ext_bump(bwg_ctx *c, const ExtScoring &sc, size_t n, const uint8_t *arena, size_t arena_len,
                   const uint64_t *q_off, const uint32_t *q_len, const uint64_t *t_off, const uint32_t *t_len,
                   const int32_t *h0, Bump *bump, const hipEvent_t ev[2], hipStream_t st, bwg_batch **out) {
    return stage_ext_impl(c, sc, n, arena, arena_len, q_off, q_len, t_off, t_len, h0, nullptr, bump, ev, st, out,
                          false);
}

int stage_side(bwg_ctx *c, const ExtScoring &sc, int tries, size_t n, const uint8_t *arena, size_t arena_len,
               const uint64_t *q_off, const uint32_t *q_len, const uint64_t *t_off, const uint32_t *t_len,
               const int32_t *h0, const int32_t *prev, hipStream_t st, bwg_batch **out, bool defer_arena) {
    bwg_batch *b = nullptr;
    int rc = stage_ext(c, sc, n, arena, arena_len, q_off, q_len, t_off, t_len, h0, nullptr, st, &b, defer_arena);
    if (rc) return rc;
    b->side_tries = tries;
    const size_t words = std::max<size_t>(n, 1) * 4;
    b->d_prev0 = (int32_t *)c->pool.alloc(words);
    b->d_prev = (int32_t *)c->pool.alloc(words);
    b->d_wout = (int32_t *)c->pool.alloc(words);
    b->d_active = (uint8_t *)c->pool.alloc(std::max<size_t>(n, 1));
    if (!b->d_prev0 || !b->d_prev || !b->d_wout || !b->d_active) rc = BWG_ENOMEM;
    if (!rc && n) rc = bwg_map_hip(hipMemcpyAsync(b->d_prev0, prev, n * 4, hipMemcpyHostToDevice, st));
    if (rc) {
        destroy_batch(c, b);
        return rc;
    }
    *out = b;
    return BWG_OK;
}

int stage_matesw(bwg_ctx *c, const KswScoring &sc, const MateRescueParams &mp, size_t n,
                 const bwg_mate_anchor *anchors, const int64_t *ma_rb, size_t n_ma, const uint8_t *arena,
                 size_t arena_len, hipStream_t st, bwg_batch **out) {
    if (hipSetDevice(c->device) != hipSuccess) return BWG_EIO;
    const size_t nj = 4 * n;
    bwg_batch *b = new_batch(c, 1, nj);
    if (!b) return BWG_ENOMEM;
    b->ksc = sc;
    b->matesw = true;
    b->n_anchors = (uint32_t)n;
    b->mp = mp;
    b->d_ref = c->d_ref2;
    b->d_contig_off = c->d_contig_off;
    b->d_contig_len = c->d_contig_len;
    b->ref_owner = c;
    c->ref_users++;
    // the mates' reverse complements go after the caller's arena
    std::vector<MateAnchor> ha(n);
    size_t rc = arena_len;
    for (size_t k = 0; k < n; ++k) {
        MateAnchor &a = ha[k];
        a.rb = anchors[k].rb;
        a.rid = anchors[k].rid;
        a.l_ms = anchors[k].l_ms;
        a.ms_off = anchors[k].ms_off;
        a.rc_off = rc;
        a.ma_begin = anchors[k].ma_begin;
        a.ma_end = anchors[k].ma_end;
        rc += anchors[k].l_ms;
    }
    // planning descriptors: the mate length and the widest window an orientation can give
    // (hi - lo + l_ms before any clamp); the device writes the actual windows before the DP runs
    b->hksw.resize(nj);
    for (size_t k = 0; k < n; ++k)
        for (int r = 0; r < 4; ++r) {
            KswJob &j = b->hksw[4 * k + r];
            const int64_t bound = mp.failed[r] ? 0 : (int64_t)mp.high[r] - mp.low[r] + anchors[k].l_ms;
            j.q_off = anchors[k].ms_off;
            j.t_off = kTFromRef;
            j.qlen = anchors[k].l_ms;
            j.tlen = (uint32_t)std::max<int64_t>(0, bound);
            j.aux0 = (uint32_t)(4 * k + r);
            j.aux1 = 0;
        }
    b->arena_len = rc;
    b->d_arena = (uint8_t *)c->pool.alloc(rc + 64);
    b->d_jobs = c->pool.alloc(std::max<size_t>(nj * sizeof(KswJob), 64));
    b->d_out = c->pool.alloc(std::max<size_t>(nj * sizeof(bwg_ksw_result), 64));
    b->d_anchors = (MateAnchor *)c->pool.alloc(std::max<size_t>(n * sizeof(MateAnchor), 64));
    b->d_ma_rb = (int64_t *)c->pool.alloc(std::max<size_t>(n_ma * 8, 64));
    b->d_win = (bwg_rescue_window *)c->pool.alloc(std::max<size_t>(nj * sizeof(bwg_rescue_window), 64));
    int rcode = BWG_OK;
    if (!b->d_arena || !b->d_jobs || !b->d_out || !b->d_anchors || !b->d_ma_rb || !b->d_win) rcode = BWG_ENOMEM;
    hipError_t e = hipSuccess;
    if (!rcode && arena_len) e = hipMemcpyAsync(b->d_arena, arena, arena_len, hipMemcpyHostToDevice, st);
    if (!rcode && e == hipSuccess && nj)
        e = hipMemcpyAsync(b->d_jobs, b->hksw.data(), nj * sizeof(KswJob), hipMemcpyHostToDevice, st);
    if (!rcode && e == hipSuccess && n)
        e = hipMemcpyAsync(b->d_anchors, ha.data(), n * sizeof(MateAnchor), hipMemcpyHostToDevice, st);
    if (!rcode && e == hipSuccess && n_ma)
        e = hipMemcpyAsync(b->d_ma_rb, ma_rb, n_ma * 8, hipMemcpyHostToDevice, st);
    if (!rcode) rcode = bwg_map_hip(e);
    if (rcode) {
        destroy_batch(c, b);
        return rcode;
    }
    *out = b;
    return BWG_OK;
}

int stage_global(bwg_ctx *c, const GlobalScoring &sc, size_t n, const uint8_t *arena, size_t arena_len,
                 const uint64_t *q_off, const uint32_t *q_len, const uint64_t *t_off, const uint32_t *t_len,
                 const int32_t *w, hipStream_t st, bwg_batch **out) {
    if (hipSetDevice(c->device) != hipSuccess) return BWG_EIO;
    bwg_batch *b = new_batch(c, 2, n);
    if (!b) return BWG_ENOMEM;
    b->gsc = sc;
    b->hglb.resize(n);
    for (size_t k = 0; k < n; ++k) {
        GlobalJob &j = b->hglb[k];
        j.q_off = q_off[k];
        j.t_off = t_off[k];
        j.qlen = q_len[k];
        j.tlen = t_len[k];
        // a band wider than both sequences changes nothing: keep w where band arithmetic cannot
        // overflow
        j.w = (int32_t)std::min<int64_t>(std::max(w[k], 0), (int64_t)q_len[k] + t_len[k] + 1);
        j.k = 0;
        j.z_off = j.ops_off = 0;
    }
    int rc = stage_buffers(c, b, arena, arena_len, nullptr, 0, n * 4, st);
    if (rc) {
        destroy_batch(c, b);
        return rc;
    }
    *out = b;
    return BWG_OK;
}

int fetch_cigar_lengths(bwg_ctx *c, bwg_batch *b, uint32_t *n_ops, hipStream_t st) {
    if (b->n == 0) return BWG_OK;
    if (!b->d_nops) return BWG_EINVAL;  // not run yet
    if (hipSetDevice(c->device) != hipSuccess) return BWG_EIO;
    hipError_t e = hipMemcpyAsync(n_ops, b->d_nops, b->n * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return bwg_map_hip(e);
    for (size_t k = 0; k < b->n; ++k)
        if (n_ops[k] == 0xffffffffu) return BWG_EINVAL;  // no alignment path inside the band
    return BWG_OK;
}

int fetch_cigars(bwg_ctx *c, bwg_batch *b, uint32_t *ops, size_t cap, hipStream_t st) {
    if (b->n == 0) return BWG_OK;
    std::vector<uint32_t> cnt(b->n);
    int rc = fetch_cigar_lengths(c, b, cnt.data(), st);
    if (rc) return rc;
    std::vector<uint64_t> off(b->n);
    uint64_t total = 0;
    for (size_t k = 0; k < b->n; ++k) {
        off[k] = total;
        total += cnt[k];
    }
    if (total > cap) return BWG_E2BIG;
    if (hipSetDevice(c->device) != hipSuccess) return BWG_EIO;
    uint64_t *d_off = (uint64_t *)c->pool.alloc(b->n * 8);
    uint32_t *d_dst = (uint32_t *)c->pool.alloc(std::max<uint64_t>(1, total) * 4);
    hipError_t e = (d_off && d_dst) ? hipSuccess : hipErrorOutOfMemory;
    if (e == hipSuccess) e = hipMemcpyAsync(d_off, off.data(), b->n * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = launch_glb_cigar_gather(st, (const GlobalJob *)glb_plan_jobs(b), (uint32_t)b->n, b->d_ops, b->d_nops,
                                    d_off, d_dst);
    if (e == hipSuccess && total) e = hipMemcpyAsync(ops, d_dst, total * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    c->pool.free(d_off);
    c->pool.free(d_dst);
    return bwg_map_hip(e);
}

int fetch_windows(bwg_ctx *c, bwg_batch *b, bwg_rescue_window *win, hipStream_t st) {
    if (b->n == 0) return BWG_OK;
    if (hipSetDevice(c->device) != hipSuccess) return BWG_EIO;
    hipError_t e = hipMemcpyAsync(win, b->d_win, b->n * sizeof(bwg_rescue_window), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return bwg_map_hip(e);
}

int fetch_side_bands(bwg_ctx *c, bwg_batch *b, int32_t *w_out, hipStream_t st) {
    if (b->n == 0) return BWG_OK;
    if (hipSetDevice(c->device) != hipSuccess) return BWG_EIO;
    hipError_t e = hipMemcpyAsync(w_out, b->d_wout, b->n * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return bwg_map_hip(e);
}

int stage_ksw(bwg_ctx *c, const KswScoring &sc, size_t n, const uint8_t *arena, size_t arena_len,
              const uint64_t *q_off, const uint32_t *q_len, const uint64_t *t_off,
              const uint32_t *t_len, const uint8_t *ref, hipStream_t st, bwg_batch **out, bool defer_arena) {
    if (hipSetDevice(c->device) != hipSuccess) return BWG_EIO;
    bwg_batch *b = new_batch(c, 1, n);
    if (!b) return BWG_ENOMEM;
    b->ksc = sc;
    b->d_ref = ref;
    b->hksw.resize(n);
    parallel_for(n, host_parts(b->tu, n), [&](size_t lo, size_t hi, int) {
        for (size_t k = lo; k < hi; ++k) {
            KswJob &j = b->hksw[k];
            j.q_off = q_off[k];
            j.t_off = t_off[k];
            j.qlen = q_len[k];
            j.tlen = t_len[k];
            j.aux0 = (uint32_t)k;
            j.aux1 = 0;
        }
    });
    int rc = stage_buffers(c, b, arena, arena_len, b->hksw.data(), n * sizeof(KswJob),
                           n * sizeof(bwg_ksw_result), st, defer_arena);
    if (rc) {
        destroy_batch(c, b);
        return rc;
    }
    *out = b;
    return BWG_OK;
}

size_t batch_device_bytes(bwg_ctx *c, int kind, size_t n, size_t arena_len, const uint32_t *q_len,
                          const uint32_t *t_len, const int32_t *h0, const ExtScoring *esc,
                          const KswScoring *ksc) {
    // stage_buffers: arena, descriptors, results
    const size_t job_bytes = n * (kind == 0 ? sizeof(ExtJob) : sizeof(KswJob));
    const size_t out_bytes = n * (kind == 0 ? sizeof(bwg_ext_result) : sizeof(bwg_ksw_result));
    size_t bytes = DevicePool::worst(arena_len + 64) + DevicePool::worst(std::max<size_t>(job_bytes, 64)) +
                   DevicePool::worst(std::max<size_t>(out_bytes, 64));
    return bytes + plan_device_bytes(c, kind, n, q_len, t_len, h0, esc, ksc);
}

// ---------------------------------------------------------------------------------------------
// plain form
// ---------------------------------------------------------------------------------------------
namespace {

int prepare_plain(bwg_ctx *c, bwg_batch *b, hipStream_t st) {
    if (b->d_plain_jobs) return BWG_OK;
    const size_t n = b->n;
    size_t total = 0;
    hipError_t e;
    if (b->kind == 0) {
        // {H,E} pairs over absolute columns 0..qlen per job
        std::vector<ExtJob> jobs(b->hext.begin(), b->hext.end());
        for (size_t k = 0; k < n; ++k) {
            jobs[k].aux = (uint32_t)total;
            total += (size_t)jobs[k].qlen + 1;
        }
        if (total >= (1ull << 32)) return BWG_E2BIG;
        b->d_plain_scratch = c->pool.alloc(total * sizeof(int2));
        b->d_plain_jobs = c->pool.alloc(n * sizeof(ExtJob));
        if (!b->d_plain_scratch || !b->d_plain_jobs) return BWG_ENOMEM;
        e = hipMemcpyAsync(b->d_plain_jobs, jobs.data(), n * sizeof(ExtJob), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        return bwg_map_hip(e);
    }
    // rescue: {H,E} pairs for the padded query (+16 for the reverse pass's own padding) and one
    // int32 row maximum per target row; slots are 16-byte aligned
    std::vector<KswJob> jobs(b->hksw.begin(), b->hksw.end());
    for (size_t k = 0; k < n; ++k) {
        const size_t bytes = ((size_t)ksw_padded((int32_t)jobs[k].qlen, b->ksc.max_sc) + 16) * 8 +
                             (size_t)jobs[k].tlen * 4;
        jobs[k].aux0 = (uint32_t)(total / 16);
        total += (bytes + 15) & ~(size_t)15;
    }
    if (total / 16 >= (1ull << 32)) return BWG_E2BIG;
    b->d_plain_scratch = c->pool.alloc(std::max<size_t>(16, total));
    b->d_plain_jobs = c->pool.alloc(n * sizeof(KswJob));
    if (!b->d_plain_scratch || !b->d_plain_jobs) return BWG_ENOMEM;
    e = hipMemcpyAsync(b->d_plain_jobs, jobs.data(), n * sizeof(KswJob), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return bwg_map_hip(e);
}

}  // namespace

// Band-retry batches keep their rounds on the device, rescue-window batches compute their windows
// there, and global alignment has only the batched kernels: those run the batched form always.
static bool runs_plain(const bwg_batch *b, int form) {
    return form == BWG_FORM_PLAIN && !b->side_tries && !b->matesw && b->kind != 2;
}

int prepare_batch(bwg_ctx *c, bwg_batch *b, int form, StreamSet &ss) {
    if (b->n == 0) return BWG_OK;
    return runs_plain(b, form) ? prepare_plain(c, b, ss.main) : plan_batched(c, b, ss.main);
}

int enqueue_batch(bwg_ctx *c, bwg_batch *b, int form, StreamSet &ss) {
    if (b->n == 0) return BWG_OK;
    const int arc = wait_arena(b);
    if (arc) return arc;
    if (runs_plain(b, form)) {
        int rc = prepare_plain(c, b, ss.main);
        if (rc) return rc;
        hipError_t e = hipEventRecord(b->ev_start, ss.main);
        if (e == hipSuccess) {
            if (b->kind == 0)
                e = launch_ext_plain(ss.main, b->d_arena, b->d_ref, (const ExtJob *)b->d_plain_jobs,
                                     (uint32_t)b->n, b->esc, (int2 *)b->d_plain_scratch,
                                     (bwg_ext_result *)b->d_out);
            else
                e = launch_ksw_plain(ss.main, b->d_arena, b->d_ref, (const KswJob *)b->d_plain_jobs,
                                     (uint32_t)b->n, b->ksc, (uint8_t *)b->d_plain_scratch,
                                     (bwg_ksw_result *)b->d_out);
        }
        if (e == hipSuccess) e = hipEventRecord(b->ev_end, ss.main);
        b->last_launches = 1;
        return bwg_map_hip(e);
    }
    int64_t launches = 0;
    int rc = enqueue_batched(c, b, ss, &launches);
    b->last_launches = launches;
    return rc;
}

void record_run(bwg_ctx *c, bwg_batch *b) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, b->ev_start, b->ev_end);
    c->last_kernel_ns = (int64_t)(ms * 1e6);
    c->last_main_ns = (int64_t)(ms * 1e6);
    c->last_main_launches = b->last_launches;
    c->launches += b->last_launches;
    c->jobs += (int64_t)b->n;
}

int run_batch(bwg_ctx *c, bwg_batch *b) {
    if (hipSetDevice(c->device) != hipSuccess) return BWG_EIO;
    std::lock_guard<std::mutex> g(c->run_mu);
    if (b->n == 0) return BWG_OK;
    b->tu = c->tuning_snapshot();
    int rc = enqueue_batch(c, b, b->tu.form, c->ss);
    if (rc) return rc;
    hipError_t e = hipStreamSynchronize(c->ss.main);
    if (e != hipSuccess) return bwg_map_hip(e);
    record_run(c, b);
    return BWG_OK;
}

int enqueue_results(bwg_ctx *c, bwg_batch *b, void *out, hipStream_t st) {
    (void)c;
    if (b->n == 0) return BWG_OK;
    const size_t bytes =
        b->n * (b->kind == 0 ? sizeof(bwg_ext_result) : b->kind == 1 ? sizeof(bwg_ksw_result) : sizeof(int32_t));
    return bwg_map_hip(hipMemcpyAsync(out, b->d_out, bytes, hipMemcpyDeviceToHost, st));
}

int fetch_results(bwg_ctx *c, bwg_batch *b, void *out) {
    if (b->n == 0) return BWG_OK;
    if (hipSetDevice(c->device) != hipSuccess) return BWG_EIO;
    std::lock_guard<std::mutex> g(c->run_mu);
    int rc = enqueue_results(c, b, out, c->ss.main);
    if (rc) return rc;
    return bwg_map_hip(hipStreamSynchronize(c->ss.main));
}

void destroy_batch(bwg_ctx *c, bwg_batch *b) {
    (void)wait_arena(b);  // a deferred upload still writing into the arena
    (void)hipSetDevice(c->device);
    c->pool.free(b->d_arena);
    c->pool.free(b->d_jobs);
    c->pool.free(b->d_out);
    c->pool.free(b->d_plain_scratch);
    c->pool.free(b->d_plain_jobs);
    c->pool.free(b->d_prev0);
    c->pool.free(b->d_prev);
    c->pool.free(b->d_wout);
    c->pool.free(b->d_active);
    c->pool.free(b->d_anchors);
    c->pool.free(b->d_ma_rb);
    c->pool.free(b->d_win);
    c->pool.free(b->d_z);
    c->pool.free(b->d_ops);
    c->pool.free(b->d_nops);
    if (b->ref_owner) b->ref_owner->ref_users--;
    release_batched(c, b);
    if (!b->borrowed_events) {
        if (b->ev_start) (void)hipEventDestroy(b->ev_start);
        if (b->ev_end) (void)hipEventDestroy(b->ev_end);
    }
    delete b;
}

}  // namespace bwg
