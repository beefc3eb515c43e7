// Serving loop: chunks of alignment work from concurrent producers, continuously batched onto the
// device under an HBM budget.
//
// Producers call bwg_server_submit from any thread; the chunk is validated there and queued.  Each
// of several slots (its own streams, pinned staging buffers and worker thread) repeatedly forms an
// epoch: it takes the queued chunks in arrival order, merging chunks of the same kind and scoring,
// up to a job cap and as far as the HBM budget admits; packs their inputs into its pinned buffer;
// stages them and enqueues the batched-form kernels and the result copy; waits for them; scatters
// the results into the producers' output arrays, records per-chunk latency (submit to results
// available) and returns the epoch's memory to the budget.  With several slots the host
// preparation of one epoch, the device work of another and the completion of a third overlap, and
// requests that arrive while epochs run join the next epoch a slot forms.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "bwg_runtime.h"

using Clock = std::chrono::steady_clock;

namespace {

struct Chunk {
    bwg_chunk c;
    bwg_ext_params ext{};
    bwg_ksw_params ksw{};
    uint64_t ticket = 0;
    Clock::time_point t_submit;
    size_t bytes = 0;
    int status = BWG_EAGAIN;  // BWG_EAGAIN while pending
    bool waited = false;      // a bwg_server_wait owns the ticket and releases it
};

struct Slot {
    bwg::StreamSet ss;
    hipEvent_t done = nullptr;
    uint8_t *pin_in = nullptr;
    size_t pin_in_cap = 0;
    uint8_t *pin_out = nullptr;
    size_t pin_out_cap = 0;
    bool busy = false;
};

struct Epoch {
    int slot = 0;
    std::vector<Chunk *> chunks[2];  // by kind
    bwg_batch *batch[2] = {nullptr, nullptr};
    size_t out_off[2] = {0, 0};      // offset of each kind's results in the slot's pinned output
    size_t bytes = 0;
    int rc = BWG_OK;
};

bool same_params(const Chunk *a, const Chunk *b) {
    if (a->c.kind != b->c.kind) return false;
    return a->c.kind == BWG_CHUNK_EXTEND ? std::memcmp(&a->ext, &b->ext, sizeof(a->ext)) == 0
                                         : std::memcmp(&a->ksw, &b->ksw, sizeof(a->ksw)) == 0;
}

}  // namespace

struct bwg_server {
    bwg_ctx *ctx = nullptr;
    uint64_t budget = 0;
    uint64_t max_epoch_jobs = 1u << 19;
    std::vector<Slot> slots;
    uint8_t *d_ref = nullptr;
    uint64_t ref_len = 0;

    std::mutex mu;
    std::condition_variable cv_dispatch, cv_complete;
    std::deque<Chunk *> pending;
    std::unordered_map<uint64_t, Chunk *> chunks;
    uint64_t next_ticket = 1;
    uint64_t in_flight_bytes = 0, peak_bytes = 0;
    bool stop = false;
    std::vector<std::thread> workers;

    // statistics
    std::vector<double> lat_ms;
    uint64_t n_chunks = 0, n_jobs = 0, n_epochs = 0;
    Clock::time_point stats_t0 = Clock::now();

    void slot_loop(int slot);
    int launch_epoch(Epoch *ep);
    bool pending_empty_hint() {
        std::lock_guard<std::mutex> g(mu);
        return pending.empty();
    }
};

namespace {

size_t chunk_bytes(bwg_ctx *ctx, const Chunk *k) {
    if (k->c.kind == BWG_CHUNK_EXTEND) {
        const bwg::ExtScoring sc = bwg::ext_scoring(k->ext);
        return bwg::batch_device_bytes(ctx, 0, k->c.n, k->c.arena_len, k->c.q_len, k->c.t_len, k->c.h0, &sc, nullptr);
    }
    const bwg::KswScoring sc = bwg::ksw_scoring(k->ksw);
    return bwg::batch_device_bytes(ctx, 1, k->c.n, k->c.arena_len, k->c.q_len, k->c.t_len, nullptr, nullptr, &sc);
}

int validate_chunk(const bwg_server *s, const bwg_chunk *c) {
    if (!c || (c->n && (!c->arena || !c->q_off || !c->q_len || !c->t_off || !c->t_len || !c->out))) return BWG_EINVAL;
    if (c->kind == BWG_CHUNK_EXTEND) {
        if (!c->ext || c->ext->m != 5 || c->ext->e_del < 1 || c->ext->e_ins < 1 || (c->n && !c->h0)) return BWG_EINVAL;
    } else if (c->kind == BWG_CHUNK_RESCUE) {
        if (!c->ksw || c->ksw->m != 5 || c->ksw->e_del < 1 || c->ksw->e_ins < 1 || c->ksw->max_sc < 1) return BWG_EINVAL;
    } else {
        return BWG_EINVAL;
    }
    const bwg::ExtScoring esc = c->kind == BWG_CHUNK_EXTEND ? bwg::ext_scoring(*c->ext) : bwg::ExtScoring{};
    const bwg::KswScoring ksc = c->kind == BWG_CHUNK_RESCUE ? bwg::ksw_scoring(*c->ksw) : bwg::KswScoring{};
    for (size_t k = 0; k < c->n; ++k) {
        if (c->q_off[k] > c->arena_len || c->q_len[k] > c->arena_len - c->q_off[k]) return BWG_EINVAL;
        if (c->kind == BWG_CHUNK_EXTEND && c->h0[k] <= 0) return BWG_EINVAL;
        if (c->kind == BWG_CHUNK_EXTEND ? !bwg::ext_job_fits(esc, c->q_len[k], c->t_len[k], c->h0[k])
                                        : !bwg::ksw_job_fits(ksc, c->q_len[k], c->t_len[k]))
            return BWG_E2BIG;
        for (uint32_t i = 0; i < c->q_len[k]; ++i)
            if (c->arena[c->q_off[k] + i] > 4) return BWG_EINVAL;
        const uint64_t to = c->t_off[k] & bwg::kTOffMask;
        if (c->t_off[k] & bwg::kTFromRef) {
            if (!s->d_ref || to > s->ref_len || c->t_len[k] > s->ref_len - to) return BWG_EINVAL;
        } else {
            if (c->t_off[k] & bwg::kTReverse) return BWG_EINVAL;
            if (to > c->arena_len || c->t_len[k] > c->arena_len - to) return BWG_EINVAL;
            for (uint32_t i = 0; i < c->t_len[k]; ++i)
                if (c->arena[to + i] > 4) return BWG_EINVAL;
        }
        if (c->kind == BWG_CHUNK_EXTEND && c->q_len[k] < 1) return BWG_EINVAL;
    }
    return BWG_OK;
}

}  // namespace

// Pack the epoch's chunks of one kind into the slot's pinned buffer and stage them as one batch.
static int stage_kind(bwg_server *s, Epoch *ep, int kind, size_t in_off) {
    auto &v = ep->chunks[kind];
    if (v.empty()) return BWG_OK;
    Slot &sl = s->slots[ep->slot];
    size_t n = 0;
    for (auto *k : v) n += k->c.n;
    std::vector<uint64_t> q_off(n), t_off(n);
    std::vector<uint32_t> q_len(n), t_len(n);
    std::vector<int32_t> h0(kind == 0 ? n : 0);
    uint8_t *arena = sl.pin_in + in_off;
    size_t a = 0, j = 0;
    for (auto *k : v) {
        const bwg_chunk &c = k->c;
        // copy the chunk's arena wholesale and rebase its offsets
        std::memcpy(arena + a, c.arena, c.arena_len);
        for (size_t r = 0; r < c.n; ++r, ++j) {
            q_off[j] = c.q_off[r] + a;
            q_len[j] = c.q_len[r];
            t_len[j] = c.t_len[r];
            t_off[j] = (c.t_off[r] & bwg::kTFromRef) ? c.t_off[r] : c.t_off[r] + a;
            if (kind == 0) h0[j] = c.h0[r];
        }
        a += c.arena_len;
    }
    int rc;
    if (kind == 0) {
        const bwg_ext_params &p = v[0]->ext;
        bwg::ExtScoring sc = bwg::ext_scoring(p);
        rc = bwg::stage_ext(s->ctx, sc, n, arena, a, q_off.data(), q_len.data(), t_off.data(), t_len.data(), h0.data(),
                            s->d_ref, sl.ss.main, &ep->batch[0]);
    } else {
        const bwg_ksw_params &p = v[0]->ksw;
        bwg::KswScoring sc = bwg::ksw_scoring(p);
        rc = bwg::stage_ksw(s->ctx, sc, n, arena, a, q_off.data(), q_len.data(), t_off.data(), t_len.data(), s->d_ref,
                            sl.ss.main, &ep->batch[1]);
    }
    return rc;
}

int bwg_server::launch_epoch(Epoch *ep) {
    Slot &sl = slots[ep->slot];
    size_t in_bytes = 0, out_bytes = 0;
    for (int kind = 0; kind < 2; ++kind)
        for (auto *k : ep->chunks[kind]) {
            in_bytes += k->c.arena_len;
            out_bytes += k->c.n * (kind == 0 ? sizeof(bwg_ext_result) : sizeof(bwg_ksw_result));
        }
    // grow the slot's pinned buffers (the slot is idle here)
    if (in_bytes + 64 > sl.pin_in_cap) {
        if (sl.pin_in) (void)hipHostFree(sl.pin_in);
        sl.pin_in_cap = std::max<size_t>(in_bytes + 64, 2 * sl.pin_in_cap);
        if (hipHostMalloc((void **)&sl.pin_in, sl.pin_in_cap, hipHostMallocDefault) != hipSuccess) {
            sl.pin_in = nullptr;
            sl.pin_in_cap = 0;
            return BWG_ENOMEM;
        }
    }
    if (out_bytes + 64 > sl.pin_out_cap) {
        if (sl.pin_out) (void)hipHostFree(sl.pin_out);
        sl.pin_out_cap = std::max<size_t>(out_bytes + 64, 2 * sl.pin_out_cap);
        if (hipHostMalloc((void **)&sl.pin_out, sl.pin_out_cap, hipHostMallocDefault) != hipSuccess) {
            sl.pin_out = nullptr;
            sl.pin_out_cap = 0;
            return BWG_ENOMEM;
        }
    }
    size_t in_off = 0, out_off = 0;
    for (int kind = 0; kind < 2; ++kind) {
        if (ep->chunks[kind].empty()) continue;
        int rc = stage_kind(this, ep, kind, in_off);
        if (rc) return rc;
        for (auto *k : ep->chunks[kind]) in_off += k->c.arena_len;
        rc = bwg::enqueue_batch(ctx, ep->batch[kind], BWG_FORM_BATCHED, sl.ss);
        if (rc) return rc;
        ep->out_off[kind] = out_off;
        rc = bwg::enqueue_results(ctx, ep->batch[kind], sl.pin_out + out_off, sl.ss.main);
        if (rc) return rc;
        out_off += ep->batch[kind]->n * (kind == 0 ? sizeof(bwg_ext_result) : sizeof(bwg_ksw_result));
    }
    return bwg_map_hip(hipEventRecord(sl.done, sl.ss.main));
}

void bwg_server::slot_loop(int si) {
    (void)hipSetDevice(ctx->device);
    Slot &sl = slots[(size_t)si];
    for (;;) {
        Epoch ep;
        ep.slot = si;
        {
            std::unique_lock<std::mutex> g(mu);
            cv_dispatch.wait(g, [&] {
                return (stop && pending.empty()) ||
                       (!pending.empty() && in_flight_bytes + pending.front()->bytes <= budget);
            });
            if (pending.empty()) return;  // stopping and drained
            // take chunks in arrival order while they merge into this epoch and fit the budget
            uint64_t jobs = 0;
            Chunk *first[2] = {nullptr, nullptr};
            while (!pending.empty()) {
                Chunk *k = pending.front();
                const int kind = k->c.kind == BWG_CHUNK_EXTEND ? 0 : 1;
                if (first[kind] && !same_params(first[kind], k)) break;
                if (jobs && jobs + k->c.n > max_epoch_jobs) break;
                if (in_flight_bytes + ep.bytes + k->bytes > budget) break;
                if (!first[kind]) first[kind] = k;
                pending.pop_front();
                ep.chunks[kind].push_back(k);
                ep.bytes += k->bytes;
                jobs += k->c.n;
            }
            in_flight_bytes += ep.bytes;
            peak_bytes = std::max(peak_bytes, in_flight_bytes);
            ++n_epochs;
        }
        if (!pending_empty_hint()) cv_dispatch.notify_one();  // let another slot start the next epoch
        int rc = launch_epoch(&ep);
        if (!rc) rc = bwg_map_hip(hipEventSynchronize(sl.done));
        else (void)hipStreamSynchronize(sl.ss.main);
        const auto now = Clock::now();
        for (int kind = 0; kind < 2; ++kind) {
            size_t off = ep.out_off[kind];
            const size_t rsz = kind == 0 ? sizeof(bwg_ext_result) : sizeof(bwg_ksw_result);
            for (auto *k : ep.chunks[kind]) {
                if (!rc && k->c.n) std::memcpy(k->c.out, sl.pin_out + off, k->c.n * rsz);
                off += k->c.n * rsz;
            }
            if (ep.batch[kind]) bwg::destroy_batch(ctx, ep.batch[kind]);
        }
        {
            std::lock_guard<std::mutex> g(mu);
            for (int kind = 0; kind < 2; ++kind)
                for (auto *k : ep.chunks[kind]) {
                    k->status = rc;
                    lat_ms.push_back(std::chrono::duration<double, std::milli>(now - k->t_submit).count());
                    ++n_chunks;
                    n_jobs += k->c.n;
                    ctx->jobs += (int64_t)k->c.n;
                }
            in_flight_bytes -= ep.bytes;
        }
        cv_complete.notify_all();
        cv_dispatch.notify_all();
    }
}

extern "C" {

int bwg_server_create(bwg_ctx *ctx, uint64_t hbm_budget_bytes, int n_slots, bwg_server **out) {
    if (!ctx || !out || n_slots < 1 || n_slots > 8 || hbm_budget_bytes == 0) return BWG_EINVAL;
    *out = nullptr;
    if (hipSetDevice(ctx->device) != hipSuccess) return BWG_EIO;
    bwg_server *s = new (std::nothrow) bwg_server();
    if (!s) return BWG_ENOMEM;
    s->ctx = ctx;
    s->budget = hbm_budget_bytes;
    s->slots.resize((size_t)n_slots);
    for (size_t k = 0; k < s->slots.size(); ++k) {
        Slot &sl = s->slots[k];
        if (sl.ss.create((int)k + 1) != BWG_OK || hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess) {
            for (auto &x : s->slots) {
                x.ss.destroy();
                if (x.done) (void)hipEventDestroy(x.done);
            }
            delete s;
            return BWG_EIO;
        }
    }
    ctx->pool.reset_peak();
    for (int k = 0; k < n_slots; ++k) s->workers.emplace_back([s, k] { s->slot_loop(k); });
    *out = s;
    return BWG_OK;
}

int bwg_server_set_reference(bwg_server *s, const uint8_t *ref, uint64_t len) {
    if (!s || (!ref && len)) return BWG_EINVAL;
    for (uint64_t k = 0; k < len; ++k)
        if (ref[k] > 4) return BWG_EINVAL;
    std::unique_lock<std::mutex> g(s->mu);
    if (!s->pending.empty() || s->in_flight_bytes) return BWG_EAGAIN;  // only while idle
    if (hipSetDevice(s->ctx->device) != hipSuccess) return BWG_EIO;
    if (s->d_ref) (void)hipFree(s->d_ref);
    s->d_ref = nullptr;
    s->ref_len = 0;
    if (!len) return BWG_OK;
    if (hipMalloc(&s->d_ref, len + 64) != hipSuccess) {
        s->d_ref = nullptr;
        return BWG_ENOMEM;
    }
    if (hipMemcpy(s->d_ref, ref, len, hipMemcpyHostToDevice) != hipSuccess) return BWG_EIO;
    s->ref_len = len;
    return BWG_OK;
}

int bwg_server_submit(bwg_server *s, const bwg_chunk *chunk, uint64_t *ticket) {
    if (!s || !chunk || !ticket) return BWG_EINVAL;
    int rc = validate_chunk(s, chunk);
    if (rc) return rc;
    Chunk *k = new (std::nothrow) Chunk();
    if (!k) return BWG_ENOMEM;
    k->c = *chunk;
    if (chunk->kind == BWG_CHUNK_EXTEND) k->ext = *chunk->ext;
    else k->ksw = *chunk->ksw;
    k->bytes = chunk_bytes(s->ctx, k);
    if (k->bytes > s->budget) {
        delete k;
        return BWG_E2BIG;
    }
    k->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->stop) {
            delete k;
            return BWG_EINVAL;
        }
        k->ticket = s->next_ticket++;
        s->chunks[k->ticket] = k;
        s->pending.push_back(k);
    }
    *ticket = k->ticket;
    s->cv_dispatch.notify_all();
    return BWG_OK;
}

int bwg_server_poll(bwg_server *s, uint64_t ticket) {
    if (!s) return BWG_EINVAL;
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->chunks.find(ticket);
    if (it == s->chunks.end()) return BWG_EINVAL;
    const int st = it->second->status;
    // a waiting thread holds the chunk: it alone releases the ticket
    if (st != BWG_EAGAIN && !it->second->waited) {
        delete it->second;
        s->chunks.erase(it);
    }
    return st;
}

int bwg_server_wait(bwg_server *s, uint64_t ticket) {
    if (!s) return BWG_EINVAL;
    std::unique_lock<std::mutex> g(s->mu);
    auto it = s->chunks.find(ticket);
    if (it == s->chunks.end()) return BWG_EINVAL;
    Chunk *k = it->second;
    if (k->waited) return BWG_EINVAL;  // one waiter per ticket
    k->waited = true;
    s->cv_complete.wait(g, [&] { return k->status != BWG_EAGAIN; });
    const int st = k->status;
    delete k;
    s->chunks.erase(ticket);
    return st;
}

int bwg_server_stats(bwg_server *s, bwg_server_counters *out, int reset) {
    if (!s || !out) return BWG_EINVAL;
    std::lock_guard<std::mutex> g(s->mu);
    std::memset(out, 0, sizeof(*out));
    out->chunks = s->n_chunks;
    out->jobs = s->n_jobs;
    out->epochs = s->n_epochs;
    out->peak_device_bytes = s->peak_bytes;
    out->peak_pool_bytes = s->ctx->pool.peak_in_use();
    out->seconds = std::chrono::duration<double>(Clock::now() - s->stats_t0).count();
    if (!s->lat_ms.empty()) {
        std::vector<double> v = s->lat_ms;
        std::sort(v.begin(), v.end());
        auto q = [&](double f) { return v[std::min(v.size() - 1, (size_t)(f * (double)(v.size() - 1) + 0.5))]; };
        out->p50_ms = q(0.50);
        out->p99_ms = q(0.99);
        out->max_ms = v.back();
        double sum = 0;
        for (double x : v) sum += x;
        out->mean_ms = sum / (double)v.size();
    }
    if (reset) {
        s->lat_ms.clear();
        s->n_chunks = s->n_jobs = s->n_epochs = 0;
        s->peak_bytes = s->in_flight_bytes;
        s->ctx->pool.reset_peak();
        s->stats_t0 = Clock::now();
    }
    return BWG_OK;
}

void bwg_server_destroy(bwg_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cv_dispatch.notify_all();
    for (auto &w : s->workers)
        if (w.joinable()) w.join();
    (void)hipSetDevice(s->ctx->device);
    for (auto &sl : s->slots) {
        sl.ss.destroy();
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.pin_in) (void)hipHostFree(sl.pin_in);
        if (sl.pin_out) (void)hipHostFree(sl.pin_out);
    }
    for (auto &kv : s->chunks) delete kv.second;
    if (s->d_ref) (void)hipFree(s->d_ref);
    delete s;
}

}  // extern "C"
