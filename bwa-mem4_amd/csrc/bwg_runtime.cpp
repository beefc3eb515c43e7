// Host runtime: device memory pool, batch staging, launch planning and result copies.
#include "bwg_runtime.h"

#include <algorithm>
#include <cstring>
#include <new>

#include "bwg_batched.h"

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
    size_t sz = round(bytes);
    std::lock_guard<std::mutex> g(mu_);
    auto it = free_.lower_bound(sz);
    if (it != free_.end() && it->first <= sz + (sz >> 1)) {
        void *p = it->second;
        size_t got = it->first;
        free_.erase(it);
        cached_ -= got;
        live_[p] = got;
        in_use_ += got;
        return p;
    }
    void *p = nullptr;
    if (hipMalloc(&p, sz) != hipSuccess) {
        // drop the cache and retry once
        for (auto &kv : free_) (void)hipFree(kv.second);
        free_.clear();
        cached_ = 0;
        (void)hipGetLastError();
        if (hipMalloc(&p, sz) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
    }
    live_[p] = sz;
    in_use_ += sz;
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
}

void DevicePool::release_all() {
    std::lock_guard<std::mutex> g(mu_);
    for (auto &kv : free_) (void)hipFree(kv.second);
    for (auto &kv : live_) (void)hipFree(kv.first);
    free_.clear();
    live_.clear();
    in_use_ = cached_ = 0;
}

// ---------------------------------------------------------------------------------------------
// staging
// ---------------------------------------------------------------------------------------------
namespace {

int stage_common(bwg_ctx *c, bwg_batch *b, const uint8_t *arena, size_t arena_len, size_t job_bytes,
                 const void *jobs_host, size_t out_bytes) {
    b->d_arena = (uint8_t *)c->pool.alloc(arena_len + 64);
    b->d_jobs = c->pool.alloc(job_bytes);
    b->d_out = c->pool.alloc(out_bytes);
    if (!b->d_arena || !b->d_jobs || !b->d_out) return BWG_ENOMEM;
    hipError_t e = hipMemcpyAsync(b->d_arena, arena, arena_len, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->d_jobs, jobs_host, job_bytes, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return bwg_map_hip(e);
}

}  // namespace

int stage_ext(bwg_ctx *c, const ExtScoring &sc, size_t n, const uint8_t *arena, size_t arena_len,
              const uint64_t *q_off, const uint32_t *q_len, const uint64_t *t_off,
              const uint32_t *t_len, const int32_t *h0, bwg_batch **out) {
    if (hipSetDevice(c->device) != hipSuccess) return BWG_EIO;
    bwg_batch *b = new (std::nothrow) bwg_batch();
    if (!b) return BWG_ENOMEM;
    b->kind = 0;
    b->n = n;
    b->esc = sc;
    b->arena_len = arena_len;
    b->q_len.assign(q_len, q_len + n);
    b->t_len.assign(t_len, t_len + n);
    b->h0.assign(h0, h0 + n);
    std::vector<ExtJob> jobs(n);
    for (size_t k = 0; k < n; ++k) {
        jobs[k].q_off = q_off[k];
        jobs[k].t_off = t_off[k];
        jobs[k].qlen = q_len[k];
        jobs[k].tlen = t_len[k];
        jobs[k].h0 = h0[k];
        jobs[k].aux = (uint32_t)k;
    }
    std::lock_guard<std::mutex> g(c->run_mu);
    int rc = stage_common(c, b, arena, arena_len, std::max<size_t>(1, n) * sizeof(ExtJob),
                          jobs.data(), std::max<size_t>(1, n) * sizeof(bwg_ext_result));
    if (rc) {
        destroy_batch(c, b);
        return rc;
    }
    *out = b;
    return BWG_OK;
}

int stage_ksw(bwg_ctx *c, const KswScoring &sc, size_t n, const uint8_t *arena, size_t arena_len,
              const uint64_t *q_off, const uint32_t *q_len, const uint64_t *t_off,
              const uint32_t *t_len, bwg_batch **out) {
    if (hipSetDevice(c->device) != hipSuccess) return BWG_EIO;
    bwg_batch *b = new (std::nothrow) bwg_batch();
    if (!b) return BWG_ENOMEM;
    b->kind = 1;
    b->n = n;
    b->ksc = sc;
    b->arena_len = arena_len;
    b->q_len.assign(q_len, q_len + n);
    b->t_len.assign(t_len, t_len + n);
    std::vector<KswJob> jobs(n);
    for (size_t k = 0; k < n; ++k) {
        jobs[k].q_off = q_off[k];
        jobs[k].t_off = t_off[k];
        jobs[k].qlen = q_len[k];
        jobs[k].tlen = t_len[k];
        jobs[k].aux0 = (uint32_t)k;
        jobs[k].aux1 = 0;
    }
    std::lock_guard<std::mutex> g(c->run_mu);
    int rc = stage_common(c, b, arena, arena_len, std::max<size_t>(1, n) * sizeof(KswJob),
                          jobs.data(), std::max<size_t>(1, n) * sizeof(bwg_ksw_result));
    if (rc) {
        destroy_batch(c, b);
        return rc;
    }
    *out = b;
    return BWG_OK;
}

// ---------------------------------------------------------------------------------------------
// plain form
// ---------------------------------------------------------------------------------------------
namespace {

int prepare_plain(bwg_ctx *c, bwg_batch *b) {
    if (b->d_plain_jobs) return BWG_OK;
    const size_t n = b->n;
    size_t total = 0;
    if (b->kind == 0) {
        // {H,E} pairs over absolute columns 0..qlen per job
        std::vector<ExtJob> jobs(n);
        hipError_t e = hipMemcpy(jobs.data(), b->d_jobs, n * sizeof(ExtJob), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return bwg_map_hip(e);
        for (size_t k = 0; k < n; ++k) {
            jobs[k].aux = (uint32_t)total;
            total += (size_t)jobs[k].qlen + 1;
        }
        if (total >= (1ull << 32)) return BWG_E2BIG;
        b->plain_scratch_bytes = total * sizeof(int2);
        b->d_plain_scratch = c->pool.alloc(b->plain_scratch_bytes);
        b->d_plain_jobs = c->pool.alloc(n * sizeof(ExtJob));
        if (!b->d_plain_scratch || !b->d_plain_jobs) return BWG_ENOMEM;
        e = hipMemcpy(b->d_plain_jobs, jobs.data(), n * sizeof(ExtJob), hipMemcpyHostToDevice);
        return bwg_map_hip(e);
    }
    // rescue: {H,E} pairs for the padded query (+16 for the reverse pass's own padding) and one
    // int32 row maximum per target row; slots are 16-byte aligned
    std::vector<KswJob> jobs(n);
    hipError_t e = hipMemcpy(jobs.data(), b->d_jobs, n * sizeof(KswJob), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return bwg_map_hip(e);
    for (size_t k = 0; k < n; ++k) {
        size_t bytes = ((size_t)ksw_padded((int32_t)jobs[k].qlen, b->ksc.max_sc) + 16) * 8 +
                       (size_t)jobs[k].tlen * 4;
        jobs[k].aux0 = (uint32_t)(total / 16);
        total += (bytes + 15) & ~(size_t)15;
    }
    if (total / 16 >= (1ull << 32)) return BWG_E2BIG;
    b->plain_scratch_bytes = total;
    b->d_plain_scratch = c->pool.alloc(std::max<size_t>(16, total));
    b->d_plain_jobs = c->pool.alloc(n * sizeof(KswJob));
    if (!b->d_plain_scratch || !b->d_plain_jobs) return BWG_ENOMEM;
    e = hipMemcpy(b->d_plain_jobs, jobs.data(), n * sizeof(KswJob), hipMemcpyHostToDevice);
    return bwg_map_hip(e);
}

}  // namespace

int run_batch(bwg_ctx *c, bwg_batch *b) {
    if (hipSetDevice(c->device) != hipSuccess) return BWG_EIO;
    std::lock_guard<std::mutex> g(c->run_mu);
    if (b->n == 0) return BWG_OK;
    const int form = c->form;
    hipError_t e = hipSuccess;
    int64_t launches = 0;
    if (form == BWG_FORM_PLAIN) {
        int rc = prepare_plain(c, b);
        if (rc) return rc;
        e = hipEventRecord(c->ev[0], c->stream);
        if (e == hipSuccess) {
            if (b->kind == 0)
                e = launch_ext_plain(c->stream, b->d_arena, (const ExtJob *)b->d_plain_jobs,
                                     (uint32_t)b->n, b->esc, (int2 *)b->d_plain_scratch,
                                     (bwg_ext_result *)b->d_out);
            else
                e = launch_ksw_plain(c->stream, b->d_arena, (const KswJob *)b->d_plain_jobs,
                                     (uint32_t)b->n, b->ksc, (uint8_t *)b->d_plain_scratch,
                                     (bwg_ksw_result *)b->d_out);
            launches = 1;
        }
        if (e == hipSuccess) e = hipEventRecord(c->ev[1], c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) return bwg_map_hip(e);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, c->ev[0], c->ev[1]);
        c->last_kernel_ns = c->last_main_ns = (int64_t)(ms * 1e6);
        c->last_main_launches = launches;
    } else {
        BatchedTiming t;
        int rc = run_batched(c, b, &t);
        if (rc) return rc;
        launches = t.launches;
        c->last_kernel_ns = t.total_ns;
        c->last_main_ns = t.main_ns;
        c->last_main_launches = t.main_launches;
    }
    c->launches += launches;
    c->jobs += (int64_t)b->n;
    return BWG_OK;
}

int fetch_results(bwg_ctx *c, bwg_batch *b, void *out) {
    if (b->n == 0) return BWG_OK;
    if (hipSetDevice(c->device) != hipSuccess) return BWG_EIO;
    size_t bytes = b->n * (b->kind == 0 ? sizeof(bwg_ext_result) : sizeof(bwg_ksw_result));
    std::lock_guard<std::mutex> g(c->run_mu);
    hipError_t e = hipMemcpyAsync(out, b->d_out, bytes, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return bwg_map_hip(e);
}

void destroy_batch(bwg_ctx *c, bwg_batch *b) {
    (void)hipSetDevice(c->device);
    c->pool.free(b->d_arena);
    c->pool.free(b->d_jobs);
    c->pool.free(b->d_out);
    c->pool.free(b->d_plain_scratch);
    c->pool.free(b->d_plain_jobs);
    release_batched(c, b);
    delete b;
}

}  // namespace bwg
