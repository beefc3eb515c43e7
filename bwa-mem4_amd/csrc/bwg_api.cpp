// Host runtime of libbwg.so: the C ABI declared in include/bwg.h.
//
// A context owns one HIP stream on one device.  A batch is staged into device memory (arena,
// job descriptors, result array, scratch), run, and its results copied back; the synchronous
// entry points bwg_extend_batch / bwg_ksw_align2_batch are exactly that sequence.  Every
// precondition the reference kernels rely on is checked on the host before anything reaches the
// device, and any HIP failure is returned as BWG_EIO: there is no CPU fallback anywhere.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "bwg_internal.h"
#include "bwg_runtime.h"

using namespace bwg;

namespace {

const char *kName = "hip-gfx950";

int check_mat(int32_t m, const int8_t *mat) {
    (void)mat;
    return m == 5 ? BWG_OK : BWG_EINVAL;
}

bool codes_ok(const uint8_t *p, size_t len) {
    // byte-wise max in blocks, so that the loop vectorises and ends early on a bad block
    size_t k = 0;
    for (; k + 4096 <= len; k += 4096) {
        uint8_t mx = 0;
        for (size_t r = 0; r < 4096; ++r) mx = std::max(mx, p[k + r]);
        if (mx > 4) return false;
    }
    uint8_t mx = 0;
    for (; k < len; ++k) mx = std::max(mx, p[k]);
    return mx <= 4;
}

// Shared validation of the arena description.  The base codes are checked over the whole arena
// first, one pass at memory speed; only an arena holding some code above 4 is checked job by job
// (such a byte may lie outside every job).  Large calls split each pass over `parts` threads.
int check_jobs(size_t n, const uint8_t *arena, size_t arena_len, const uint64_t *q_off,
               const uint32_t *q_len, const uint64_t *t_off, const uint32_t *t_len, int parts) {
    if (n == 0) return BWG_OK;
    if (!arena || !q_off || !q_len || !t_off || !t_len) return BWG_EINVAL;
    std::atomic<bool> bad{false};
    parallel_for(n, parts, [&](size_t b, size_t e, int) {
        for (size_t k = b; k < e; ++k)
            if (q_off[k] > arena_len || q_len[k] > arena_len - q_off[k] || t_off[k] > arena_len ||
                t_len[k] > arena_len - t_off[k]) {
                bad = true;
                return;
            }
    });
    if (bad) return BWG_EINVAL;
    const int arena_parts = (int)std::max<size_t>(1, std::min<size_t>((size_t)parts, arena_len >> 20));
    parallel_for(arena_len, arena_parts, [&](size_t b, size_t e, int) {
        if (!codes_ok(arena + b, e - b)) bad = true;
    });
    if (!bad) return BWG_OK;
    bad = false;
    parallel_for(n, parts, [&](size_t b, size_t e, int) {
        for (size_t k = b; k < e; ++k)
            if (!codes_ok(arena + q_off[k], q_len[k]) || !codes_ok(arena + t_off[k], t_len[k])) {
                bad = true;
                return;
            }
    });
    return bad ? BWG_EINVAL : BWG_OK;
}

// The first job (in index order) for which test(k) returns an error code, tested over `parts`
// contiguous ranges: the lowest range with a failure holds it.
template <class Test>
int first_failure(size_t n, int parts, Test test) {
    std::vector<int> code((size_t)std::max(parts, 1), BWG_OK);
    parallel_for(n, parts, [&](size_t b, size_t e, int part) {
        for (size_t k = b; k < e; ++k) {
            const int r = test(k);
            if (r) {
                code[(size_t)part] = r;
                return;
            }
        }
    });
    for (int r : code)
        if (r) return r;
    return BWG_OK;
}

}  // namespace

int bwg_map_hip(hipError_t e) {
    if (e == hipSuccess) return BWG_OK;
    if (e == hipErrorOutOfMemory) return BWG_ENOMEM;
    return BWG_EIO;
}

extern "C" {

const char *bwg_strerror(int code) {
    switch (code) {
    case BWG_OK: return "ok";
    case BWG_EIO: return "device error";
    case BWG_E2BIG: return "job exceeds a supported size";
    case BWG_EAGAIN: return "HBM budget cannot admit the chunk now";
    case BWG_ENOMEM: return "out of memory";
    case BWG_ENODEV: return "no usable gfx950 device";
    case BWG_EINVAL: return "invalid argument (a kernel precondition is violated)";
    default: return "unknown error";
    }
}

int bwg_create(int device, bwg_ctx **out) {
    if (!out) return BWG_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return BWG_ENODEV;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return BWG_ENODEV;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return BWG_ENODEV;
    bwg_ctx *c = new (std::nothrow) bwg_ctx();
    if (!c) return BWG_ENOMEM;
    c->device = device;
    c->num_cus = prop.multiProcessorCount;
    for (float &v : c->dp_share) v = -1.0f;
    // the allocation-free entry point's scratch address is looked up here, with the device current
    if (hipSetDevice(device) != hipSuccess || c->ss.create() != BWG_OK ||
        hipEventCreate(&c->na_ev[0]) != hipSuccess || hipEventCreate(&c->na_ev[1]) != hipSuccess ||
        noalloc_setup(device) != BWG_OK) {
        bwg_destroy(c);
        return BWG_EIO;
    }
    *out = c;
    return BWG_OK;
}

void bwg_destroy(bwg_ctx *c) {
    if (!c) return;
    if (c->coalescer) bwg_server_destroy(c->coalescer);
    (void)hipSetDevice(c->device);
    c->ss.destroy();
    for (auto &l : c->lanes) l->destroy();
    c->lanes.clear();
    c->free_lanes.clear();
    destroy_uploads(c);
    for (hipEvent_t ev : c->na_ev)
        if (ev) (void)hipEventDestroy(ev);
    c->pool.release_all();
    if (c->d_ref2) (void)hipFree(c->d_ref2);
    if (c->d_contig_off) (void)hipFree(c->d_contig_off);
    if (c->d_contig_len) (void)hipFree(c->d_contig_len);
    delete c;
}

const char *bwg_name(const bwg_ctx *) { return kName; }

int bwg_set_option(bwg_ctx *c, int key, int64_t v) {
    if (!c) return BWG_EINVAL;
    std::lock_guard<std::mutex> g(c->opt_mu);
    Tuning &t = c->tuning;
    switch (key) {
    case BWG_OPT_FORM:
        if (v != BWG_FORM_PLAIN && v != BWG_FORM_BATCHED) return BWG_EINVAL;
        t.form = (int)v;
        return BWG_OK;
    case BWG_OPT_CONCURRENT_BINS:
        t.concurrent_bins = v != 0;
        return BWG_OK;
    case BWG_OPT_GRAPHS:
        t.use_graphs = v != 0;
        return BWG_OK;
    case BWG_OPT_EXT_THREAD_MAX_Q:
        if (v < 0 || v > 32) return BWG_EINVAL;
        t.ext_thread_max_q16 = (int)std::min<int64_t>(v, 16);
        t.ext_thread_max_q32 = (int)v;
        return BWG_OK;
    case BWG_OPT_AUTOTUNE:
        t.autotune = v != 0;
        c->tuner.clear();
        return BWG_OK;
    case BWG_OPT_EXT_K_STEP:
        t.ext_k_step = v != 0;
        return BWG_OK;
    case BWG_OPT_KSW_K_STEP:
        t.ksw_k_step = v != 0;
        return BWG_OK;
    case BWG_OPT_UNGAPPED:
        t.ungapped = v != 0;
        return BWG_OK;
    case BWG_OPT_EXT_SCORE_WIDTH:
        if (v != 16 && v != 32) return BWG_EINVAL;
        t.ext_force_wide = v == 32;
        c->tuner.clear();
        return BWG_OK;
    case BWG_OPT_KSW_SCORE_WIDTH:
        if (v != 16 && v != 32) return BWG_EINVAL;
        t.ksw_force_wide = v == 32;
        c->tuner.clear();
        return BWG_OK;
    case BWG_OPT_LANES: {
        if (v < 1 || v > 16) return BWG_EINVAL;
        std::lock_guard<std::mutex> lg(c->lane_mu);
        c->max_lanes = (int)v;  // lanes already created stay; no more than v are handed out anew
        return BWG_OK;
    }
    case BWG_OPT_POOL_CACHE_MB:
        if (v < 0) return BWG_EINVAL;
        if (hipSetDevice(c->device) != hipSuccess) return BWG_EIO;
        c->pool.set_cache_cap((size_t)v << 20);
        return BWG_OK;
    case BWG_OPT_POOL_PEAK_RESET:
        c->pool.reset_peak();
        return BWG_OK;
    case BWG_OPT_ITEM_TIMING:
        t.item_timing = v != 0;
        return BWG_OK;
    case BWG_OPT_COALESCE:
        t.coalesce = v != 0;
        return BWG_OK;
    case BWG_OPT_HOST_THREADS:
        if (v < 0 || v > 16) return BWG_EINVAL;
        t.host_threads = (int)v;
        return BWG_OK;
    case BWG_OPT_EXT_GROUP_LANES:
        if (v != 8 && v != 16) return BWG_EINVAL;
        t.ext_group_lanes = (int)v;
        return BWG_OK;
    case BWG_OPT_KSW_GROUP_LANES:
        if (v != 8 && v != 16) return BWG_EINVAL;
        t.ksw_group_lanes = (int)v;
        return BWG_OK;
    case BWG_OPT_BIN_STREAMS:
        if (v < 1 || v > kAuxStreams + 1) return BWG_EINVAL;
        t.bin_streams = (int)v;
        return BWG_OK;
    default: return BWG_EINVAL;
    }
}

int bwg_tune_report(const bwg_ctx *c, char *buf, size_t len) {
    if (!c) return BWG_EINVAL;
    const std::string r = c->tuner.report();
    if (buf && len) {
        const size_t m = std::min(len - 1, r.size());
        std::memcpy(buf, r.data(), m);
        buf[m] = 0;
    }
    return (int)r.size();
}

int64_t bwg_get_stat(const bwg_ctx *c, int key) {
    if (!c) return -1;
    const Tuning t = c->tuning_snapshot();
    switch (key) {
    case BWG_OPT_FORM: return t.form;
    case BWG_OPT_CONCURRENT_BINS: return t.concurrent_bins;
    case BWG_OPT_GRAPHS: return t.use_graphs;
    case BWG_OPT_EXT_THREAD_MAX_Q: return t.ext_thread_max_q32;
    case BWG_OPT_AUTOTUNE: return t.autotune;
    case BWG_OPT_EXT_K_STEP: return t.ext_k_step;
    case BWG_OPT_KSW_K_STEP: return t.ksw_k_step;
    case BWG_OPT_EXT_SCORE_WIDTH: return t.ext_force_wide ? 32 : 16;
    case BWG_OPT_UNGAPPED: return t.ungapped;
    case BWG_OPT_KSW_SCORE_WIDTH: return t.ksw_force_wide ? 32 : 16;
    case BWG_OPT_LANES: return c->max_lanes;
    case BWG_OPT_POOL_CACHE_MB: return (int64_t)(c->pool.cache_cap() >> 20);
    case BWG_OPT_ITEM_TIMING: return t.item_timing;
    case BWG_OPT_COALESCE: return t.coalesce;
    case BWG_OPT_HOST_THREADS: return t.host_threads;
    case BWG_OPT_EXT_GROUP_LANES: return t.ext_group_lanes;
    case BWG_OPT_KSW_GROUP_LANES: return t.ksw_group_lanes;
    case BWG_OPT_BIN_STREAMS: return t.bin_streams;
    case BWG_STAT_LAUNCHES: return c->launches.load();
    case BWG_STAT_JOBS: return c->jobs.load();
    case BWG_STAT_LAST_KERNEL_NS: return c->last_kernel_ns.load();
    case BWG_STAT_LAST_MAIN_NS: return c->last_main_ns.load();
    case BWG_STAT_LAST_MAIN_LAUNCHES: return c->last_main_launches.load();
    case BWG_STAT_POOL_IN_USE: return (int64_t)c->pool.bytes_in_use();
    case BWG_STAT_POOL_CACHED: return (int64_t)c->pool.bytes_cached();
    case BWG_STAT_POOL_PEAK: return (int64_t)c->pool.peak_in_use();
    case BWG_STAT_PHASE_VALIDATE_NS: return c->phase_ns[kPhValidate].load();
    case BWG_STAT_PHASE_STAGE_NS: return c->phase_ns[kPhStage].load();
    case BWG_STAT_PHASE_PLAN_NS: return c->phase_ns[kPhPlan].load();
    case BWG_STAT_PHASE_KERNELS_NS: return c->phase_ns[kPhKernels].load();
    case BWG_STAT_PHASE_FETCH_NS: return c->phase_ns[kPhFetch].load();
    case BWG_STAT_PHASE_TOTAL_NS: return c->phase_ns[kPhTotal].load();
    default: return -1;
    }
}

int bwg_hbm_copy(bwg_ctx *c, size_t bytes, int reps, double *gb_per_s) {
    if (!c || bytes < 16 || bytes % 16 || reps < 1) return BWG_EINVAL;
    if (hipSetDevice(c->device) != hipSuccess) return BWG_EIO;
    std::lock_guard<std::mutex> g(c->run_mu);
    void *src = nullptr, *dst = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&src, bytes);
    if (e == hipSuccess) e = hipMalloc(&dst, bytes);
    if (e == hipSuccess) e = hipMemsetAsync(src, 1, bytes, c->ss.main);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    // 8 workgroups per CU keep enough 16-byte loads in flight to saturate HBM
    const uint32_t blocks = (uint32_t)std::max(1, c->num_cus) * 8;
    if (e == hipSuccess) e = launch_hbm_copy(c->ss.main, src, dst, bytes, blocks);  // warm-up
    if (e == hipSuccess) e = hipEventRecord(e0, c->ss.main);
    for (int r = 0; r < reps && e == hipSuccess; ++r) e = launch_hbm_copy(c->ss.main, src, dst, bytes, blocks);
    if (e == hipSuccess) e = hipEventRecord(e1, c->ss.main);
    if (e == hipSuccess) e = hipStreamSynchronize(c->ss.main);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e == hipSuccess && gb_per_s) *gb_per_s = 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9;
    c->launches += reps + 1;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (src) (void)hipFree(src);
    if (dst) (void)hipFree(dst);
    return bwg_map_hip(e);
}

}  // extern "C"

namespace {

using Clock = std::chrono::steady_clock;
inline int64_t ns_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(b - a).count();
}

// Preconditions of ksw_extend2 (crates/bwa-extend/src/lib.rs:181-195, sw.rs:203-230) and the size
// limits of include/bwg.h.
int check_ext(const bwg_ext_params *p, size_t n, const uint8_t *arena, size_t arena_len, const uint64_t *q_off,
              const uint32_t *q_len, const uint64_t *t_off, const uint32_t *t_len, const int32_t *h0,
              int parts = 1) {
    int rc = check_mat(p->m, p->mat);
    if (rc) return rc;
    if (p->e_del < 1 || p->e_ins < 1) return BWG_EINVAL;
    rc = check_jobs(n, arena, arena_len, q_off, q_len, t_off, t_len, parts);
    if (rc) return rc;
    if (n && !h0) return BWG_EINVAL;
    const ExtScoring sc = ext_scoring(*p);
    return first_failure(n, parts, [&](size_t k) {
        if (h0[k] <= 0 || q_len[k] < 1) return (int)BWG_EINVAL;
        return ext_job_fits(sc, q_len[k], t_len[k], h0[k]) ? (int)BWG_OK : (int)BWG_E2BIG;
    });
}

// Preconditions of ksw_align2 / batched_ksw_align2 (crates/bwa-extend/src/sw.rs:1108-1120,
// crates/bwa-neon/src/matesw.rs:299-309) and the size limits.
int check_ksw(const bwg_ksw_params *p, size_t n, const uint8_t *arena, size_t arena_len, const uint64_t *q_off,
              const uint32_t *q_len, const uint64_t *t_off, const uint32_t *t_len, int parts = 1) {
    int rc = check_mat(p->m, p->mat);
    if (rc) return rc;
    if (p->e_del < 1 || p->e_ins < 1 || p->max_sc < 1) return BWG_EINVAL;
    rc = check_jobs(n, arena, arena_len, q_off, q_len, t_off, t_len, parts);
    if (rc) return rc;
    const KswScoring sc = ksw_scoring(*p);
    return first_failure(n, parts, [&](size_t k) {
        return ksw_job_fits(sc, q_len[k], t_len[k]) ? (int)BWG_OK : (int)BWG_E2BIG;
    });
}

int check_global(const bwg_global_params *p, size_t n, const uint8_t *arena, size_t arena_len, const uint64_t *q_off,
                 const uint32_t *q_len, const uint64_t *t_off, const uint32_t *t_len, const int32_t *w) {
    if (n && !w) return BWG_EINVAL;
    int rc = check_mat(p->m, p->mat);
    if (rc) return rc;
    if (p->e_del < 1 || p->e_ins < 1 || p->o_del < 0 || p->o_ins < 0) return BWG_EINVAL;
    rc = check_jobs(n, arena, arena_len, q_off, q_len, t_off, t_len, 1);
    if (rc) return rc;
    for (size_t k = 0; k < n; ++k) {
        if (q_len[k] > kMaxJobLen || t_len[k] > kMaxJobLen) return BWG_E2BIG;
        // a query past the group kernel's columns runs one job per thread with a code byte per band
        // cell: at most 1 GiB of them a job
        const uint64_t ncol = std::min<uint64_t>(q_len[k], 2 * (uint64_t)std::max(w[k], 0) + 1);
        if (q_len[k] > kGlbGroupMaxQ && (uint64_t)t_len[k] * ncol > (1ull << 30)) return BWG_E2BIG;
        // the band must reach the last cell (cigar.rs:441-443 guarantees it)
        if ((int64_t)w[k] < std::abs((int64_t)q_len[k] - (int64_t)t_len[k])) return BWG_EINVAL;
        // gap costs of a whole side stay far from MINUS_INF
        const int64_t span = (int64_t)std::max(q_len[k], t_len[k]) + 1;
        if (p->o_del + (int64_t)p->e_del * span >= (1 << 28) || p->o_ins + (int64_t)p->e_ins * span >= (1 << 28))
            return BWG_E2BIG;
    }
    return BWG_OK;
}

GlobalScoring global_scoring(const bwg_global_params &p) {
    GlobalScoring sc{};
    for (int k = 0; k < 25; ++k) sc.mat[k] = p.mat[k];
    sc.o_del = p.o_del;
    sc.e_del = p.e_del;
    sc.o_ins = p.o_ins;
    sc.e_ins = p.e_ins;
    return sc;
}

int check_matesw(bwg_ctx *c, const bwg_ksw_params *p, int32_t min_seed_len, const bwg_pe_orient pes[4], size_t n,
                 const bwg_mate_anchor *anchors, size_t n_ma, const uint8_t *arena, size_t arena_len,
                 MateRescueParams *mp) {
    int rc = check_mat(p->m, p->mat);
    if (rc) return rc;
    if (p->e_del < 1 || p->e_ins < 1 || p->max_sc < 1) return BWG_EINVAL;
    if (!c->d_ref2) return BWG_EINVAL;  // bwg_set_reference first
    if (n > (size_t)(INT32_MAX / 4)) return BWG_E2BIG;
    *mp = MateRescueParams{};
    mp->l_pac = c->l_pac;
    mp->min_seed_len = min_seed_len;
    mp->n_contigs = c->n_contigs;
    for (int r = 0; r < 4; ++r) {
        mp->low[r] = pes[r].low;
        mp->high[r] = pes[r].high;
        mp->failed[r] = pes[r].failed ? 1 : 0;
    }
    const KswScoring sc = ksw_scoring(*p);
    for (size_t k = 0; k < n; ++k) {
        const bwg_mate_anchor &a = anchors[k];
        if (a.ms_off > arena_len || a.l_ms > arena_len - a.ms_off || !codes_ok(arena + a.ms_off, a.l_ms))
            return BWG_EINVAL;
        if (a.ma_begin > a.ma_end || a.ma_end > n_ma) return BWG_EINVAL;
        for (int r = 0; r < 4; ++r) {
            if (mp->failed[r]) continue;
            const int64_t bound = (int64_t)mp->high[r] - mp->low[r] + a.l_ms;
            if (bound > (int64_t)kMaxJobLen || !ksw_job_fits(sc, a.l_ms, (uint32_t)std::max<int64_t>(bound, 0)))
                return BWG_E2BIG;
        }
    }
    return BWG_OK;
}

// Concurrent small synchronous calls.  A call counts itself in c->sync_active for its duration; a
// call of up to kCoalesceJobs jobs that finds another one running is submitted as a chunk to the
// context's own server, which merges it with whatever else is waiting into one device batch.
constexpr size_t kCoalesceJobs = 65536;

struct ActiveCall {
    bwg_ctx *c;
    int before;
    explicit ActiveCall(bwg_ctx *ctx) : c(ctx), before(ctx->sync_active.fetch_add(1)) {}
    ~ActiveCall() { c->sync_active.fetch_sub(1); }
};

bool coalesces(bwg_ctx *c, const ActiveCall &a, size_t n) {
    return n <= kCoalesceJobs && a.before > 0 && c->tuning_snapshot().coalesce;
}

int run_coalesced(bwg_ctx *c, const bwg_chunk &ch) {
    bwg_server *s;
    {
        std::lock_guard<std::mutex> g(c->coalesce_mu);
        if (!c->coalescer) {
            const int rc = bwg_server_create(c, 16ull << 30, 2, &c->coalescer);
            if (rc) return rc;
        }
        s = c->coalescer;
    }
    uint64_t ticket = 0;
    const int rc = bwg_server_submit(s, &ch, &ticket);
    return rc ? rc : bwg_server_wait(s, ticket);
}

// Staged-batch calls: stage on the context's own streams under run_mu, then wait, so that the
// caller may free its buffers on return.
template <class Stage>
int stage_on_context(bwg_ctx *c, Stage stage) {
    if (hipSetDevice(c->device) != hipSuccess) return BWG_EIO;
    std::lock_guard<std::mutex> g(c->run_mu);
    int rc = stage(c->ss.main);
    if (!rc) rc = bwg_map_hip(hipStreamSynchronize(c->ss.main));
    return rc;
}

// One synchronous call on an execution lane: stage, plan, enqueue the kernels and the result
// copies, wait, and record the phases (validation time measured by the caller).  `stage` makes the
// batch on the lane's main stream; `fetch` enqueues (or performs) the copies of its outputs on it.
template <class Stage, class Fetch>
int run_on_lane(bwg_ctx *c, Clock::time_point t_begin, Clock::time_point t_valid, Stage stage, Fetch fetch) {
    if (hipSetDevice(c->device) != hipSuccess) return BWG_EIO;
    ExecLane *L = acquire_lane(c);
    if (!L) return BWG_ENOMEM;
    bwg_batch *b = nullptr;
    int rc = stage(L->ss.main, &b);
    const Clock::time_point t_staged = Clock::now();
    if (!rc) rc = prepare_batch(c, b, b->tu.form, L->ss);
    const Clock::time_point t_planned = Clock::now();
    if (!rc) rc = wait_arena(b);  // the rest of an upload the planning did not hide
    const Clock::time_point t_uploaded = Clock::now();
    if (!rc) rc = enqueue_batch(c, b, b->tu.form, L->ss);
    if (!rc) rc = bwg_map_hip(hipStreamSynchronize(L->ss.main));
    const Clock::time_point t_run = Clock::now();
    if (!rc) {
        record_run(c, b);
        rc = fetch(b, L->ss.main);
    }
    if (!rc) rc = bwg_map_hip(hipStreamSynchronize(L->ss.main));
    const Clock::time_point t_end = Clock::now();
    if (b) destroy_batch(c, b);
    release_lane(c, L);
    if (!rc) {
        c->phase_ns[kPhValidate] = ns_between(t_begin, t_valid);
        c->phase_ns[kPhStage] = ns_between(t_valid, t_staged) + ns_between(t_planned, t_uploaded);
        c->phase_ns[kPhPlan] = ns_between(t_staged, t_planned);
        c->phase_ns[kPhKernels] = ns_between(t_uploaded, t_run);
        c->phase_ns[kPhFetch] = ns_between(t_run, t_end);
        c->phase_ns[kPhTotal] = ns_between(t_begin, t_end);
    }
    return rc;
}

}  // namespace

extern "C" {

// ---- staged batches -------------------------------------------------------------------------

int bwg_batch_stage_ext(bwg_ctx *c, const bwg_ext_params *p, size_t n, const uint8_t *arena,
                        size_t arena_len, const uint64_t *q_off, const uint32_t *q_len,
                        const uint64_t *t_off, const uint32_t *t_len, const int32_t *h0,
                        bwg_batch **out) {
    if (!c || !p || !out) return BWG_EINVAL;
    *out = nullptr;
    int rc = check_ext(p, n, arena, arena_len, q_off, q_len, t_off, t_len, h0, host_parts(c->tuning_snapshot(), n));
    if (rc) return rc;
    const ExtScoring sc = ext_scoring(*p);
    return stage_on_context(c, [&](hipStream_t st) {
        return stage_ext(c, sc, n, arena, arena_len, q_off, q_len, t_off, t_len, h0, nullptr, st, out);
    });
}

int bwg_batch_stage_side(bwg_ctx *c, const bwg_ext_params *p, int max_band_try, size_t n, const uint8_t *arena,
                         size_t arena_len, const uint64_t *q_off, const uint32_t *q_len, const uint64_t *t_off,
                         const uint32_t *t_len, const int32_t *h0, const int32_t *prev, bwg_batch **out) {
    if (!c || !p || !out) return BWG_EINVAL;
    *out = nullptr;
    if (max_band_try < 1 || max_band_try > 8 || p->w < 0) return BWG_EINVAL;
    if (p->w > (INT32_MAX >> (max_band_try - 1))) return BWG_E2BIG;
    int rc = check_ext(p, n, arena, arena_len, q_off, q_len, t_off, t_len, h0, host_parts(c->tuning_snapshot(), n));
    if (rc) return rc;
    if (n && !prev) return BWG_EINVAL;
    const ExtScoring sc = ext_scoring(*p);
    return stage_on_context(c, [&](hipStream_t st) {
        return stage_side(c, sc, max_band_try, n, arena, arena_len, q_off, q_len, t_off, t_len, h0, prev, st, out);
    });
}

int bwg_batch_side_bands(bwg_ctx *c, bwg_batch *b, int32_t *w_out) {
    if (!c || !b || (!w_out && b->n) || !b->side_tries) return BWG_EINVAL;
    std::lock_guard<std::mutex> g(c->run_mu);
    return fetch_side_bands(c, b, w_out, c->ss.main);
}

int bwg_batch_stage_global(bwg_ctx *c, const bwg_global_params *p, size_t n, const uint8_t *arena, size_t arena_len,
                           const uint64_t *q_off, const uint32_t *q_len, const uint64_t *t_off, const uint32_t *t_len,
                           const int32_t *w, bwg_batch **out) {
    if (!c || !p || !out) return BWG_EINVAL;
    *out = nullptr;
    int rc = check_global(p, n, arena, arena_len, q_off, q_len, t_off, t_len, w);
    if (rc) return rc;
    const GlobalScoring sc = global_scoring(*p);
    return stage_on_context(c, [&](hipStream_t st) {
        return stage_global(c, sc, n, arena, arena_len, q_off, q_len, t_off, t_len, w, st, out);
    });
}

int bwg_batch_cigar_lengths(bwg_ctx *c, bwg_batch *b, uint32_t *n_cigar) {
    if (!c || !b || b->kind != 2 || (!n_cigar && b->n)) return BWG_EINVAL;
    std::lock_guard<std::mutex> g(c->run_mu);
    return fetch_cigar_lengths(c, b, n_cigar, c->ss.main);
}

int bwg_batch_cigars(bwg_ctx *c, bwg_batch *b, uint32_t *cigar, size_t cap) {
    if (!c || !b || b->kind != 2 || (!cigar && cap)) return BWG_EINVAL;
    std::lock_guard<std::mutex> g(c->run_mu);
    return fetch_cigars(c, b, cigar, cap, c->ss.main);
}

int bwg_batch_stage_matesw(bwg_ctx *c, const bwg_ksw_params *p, int32_t min_seed_len, const bwg_pe_orient pes[4],
                           size_t n, const bwg_mate_anchor *anchors, const int64_t *ma_rb, size_t n_ma,
                           const uint8_t *arena, size_t arena_len, bwg_batch **out) {
    if (!c || !p || !pes || !out || (n && !anchors) || (n_ma && !ma_rb)) return BWG_EINVAL;
    *out = nullptr;
    MateRescueParams mp;
    int rc = check_matesw(c, p, min_seed_len, pes, n, anchors, n_ma, arena, arena_len, &mp);
    if (rc) return rc;
    const KswScoring sc = ksw_scoring(*p);
    return stage_on_context(c, [&](hipStream_t st) {
        return stage_matesw(c, sc, mp, n, anchors, ma_rb, n_ma, arena, arena_len, st, out);
    });
}

int bwg_batch_windows(bwg_ctx *c, bwg_batch *b, bwg_rescue_window *win) {
    if (!c || !b || !b->matesw || (!win && b->n)) return BWG_EINVAL;
    std::lock_guard<std::mutex> g(c->run_mu);
    return fetch_windows(c, b, win, c->ss.main);
}

int bwg_batch_stage_ksw(bwg_ctx *c, const bwg_ksw_params *p, size_t n, const uint8_t *arena,
                        size_t arena_len, const uint64_t *q_off, const uint32_t *q_len,
                        const uint64_t *t_off, const uint32_t *t_len, bwg_batch **out) {
    if (!c || !p || !out) return BWG_EINVAL;
    *out = nullptr;
    int rc = check_ksw(p, n, arena, arena_len, q_off, q_len, t_off, t_len, host_parts(c->tuning_snapshot(), n));
    if (rc) return rc;
    const KswScoring sc = ksw_scoring(*p);
    return stage_on_context(c, [&](hipStream_t st) {
        return stage_ksw(c, sc, n, arena, arena_len, q_off, q_len, t_off, t_len, nullptr, st, out);
    });
}

int bwg_batch_ungapped_hits(bwg_ctx *c, bwg_batch *b, uint64_t *hits) {
    if (!c || !b || !hits) return BWG_EINVAL;
    if (hipSetDevice(c->device) != hipSuccess) return BWG_EIO;
    std::lock_guard<std::mutex> g(c->run_mu);
    return batched_ungapped_hits(c, b, hits);
}

int bwg_batch_report(bwg_ctx *c, bwg_batch *b, char *buf, size_t len) {
    if (!c || !b) return BWG_EINVAL;
    if (hipSetDevice(c->device) != hipSuccess) return BWG_EIO;
    std::lock_guard<std::mutex> g(c->run_mu);
    const std::string r = batched_report(b);
    if (buf && len) {
        const size_t m = std::min(len - 1, r.size());
        std::memcpy(buf, r.data(), m);
        buf[m] = 0;
    }
    return (int)r.size();
}

int bwg_batch_job_bins(bwg_ctx *c, bwg_batch *b, int32_t *bin) {
    if (!c || !b || (!bin && b->n)) return BWG_EINVAL;
    std::lock_guard<std::mutex> g(c->run_mu);
    return b->n ? batched_job_bins(b, bin) : BWG_OK;
}

int bwg_batch_run(bwg_ctx *c, bwg_batch *b) {
    if (!c || !b) return BWG_EINVAL;
    return run_batch(c, b);
}

int bwg_batch_results(bwg_ctx *c, bwg_batch *b, void *out) {
    if (!c || !b || (!out && b->n)) return BWG_EINVAL;
    return fetch_results(c, b, out);
}

void bwg_batch_destroy(bwg_ctx *c, bwg_batch *b) {
    if (!c || !b) return;
    destroy_batch(c, b);
}

// ---- synchronous entry points (execution lanes) ---------------------------------------------

int bwg_extend_batch(bwg_ctx *c, const bwg_ext_params *p, size_t n, const uint8_t *arena,
                     size_t arena_len, const uint64_t *q_off, const uint32_t *q_len,
                     const uint64_t *t_off, const uint32_t *t_len, const int32_t *h0,
                     bwg_ext_result *out) {
    if (!c || !p) return BWG_EINVAL;
    if (n == 0) return BWG_OK;
    if (!out) return BWG_EINVAL;
    const Clock::time_point t0 = Clock::now();
    ActiveCall active(c);
    if (coalesces(c, active, n)) {
        bwg_chunk ch{};
        ch.kind = BWG_CHUNK_EXTEND;
        ch.ext = p;
        ch.n = n;
        ch.arena = arena;
        ch.arena_len = arena_len;
        ch.q_off = q_off;
        ch.q_len = q_len;
        ch.t_off = t_off;
        ch.t_len = t_len;
        ch.h0 = h0;
        ch.out = out;
        return run_coalesced(c, ch);  // the server validates the chunk as this call would
    }
    int rc = check_ext(p, n, arena, arena_len, q_off, q_len, t_off, t_len, h0, host_parts(c->tuning_snapshot(), n));
    if (rc) return rc;
    const ExtScoring sc = ext_scoring(*p);
    return run_on_lane(
        c, t0, Clock::now(),
        [&](hipStream_t st, bwg_batch **b) {
            return stage_ext(c, sc, n, arena, arena_len, q_off, q_len, t_off, t_len, h0, nullptr, st, b, true);
        },
        [&](bwg_batch *b, hipStream_t st) { return enqueue_results(c, b, out, st); });
}

int bwg_extend_batch_noalloc(bwg_ctx *c, const bwg_ext_params *p, size_t n, const uint8_t *arena,
                             size_t arena_len, const uint64_t *q_off, const uint32_t *q_len,
                             const uint64_t *t_off, const uint32_t *t_len, const int32_t *h0,
                             bwg_ext_result *out) {
    if (!c || !p) return BWG_EINVAL;
    if (n == 0) return BWG_OK;
    if (!out) return BWG_EINVAL;
    int rc = check_ext(p, n, arena, arena_len, q_off, q_len, t_off, t_len, h0, host_parts(c->tuning_snapshot(), n));
    if (rc) return rc;
    const ExtScoring sc = ext_scoring(*p);
    uint8_t *base = noalloc_scratch(c->device);
    if (!base) return BWG_EIO;
    // everything the call places in the scratch: the staged arena, descriptors and results
    // (stage_buffers) and the plan's arrays (plan_ext); it never tunes
    const size_t need = bump_round(arena_len + 64) + bump_round(std::max<size_t>(n * sizeof(ExtJob), 64)) +
                        bump_round(std::max<size_t>(n * sizeof(bwg_ext_result), 64)) +
                        plan_device_bytes(c, 0, n, q_len, t_len, h0, &sc, nullptr, true);
    if (need > kNoAllocScratchBytes) return BWG_E2BIG;
    if (hipSetDevice(c->device) != hipSuccess) return BWG_EIO;
    std::lock_guard<std::mutex> g_dev(noalloc_call_mutex(c->device));  // one call per device: shared array
    std::lock_guard<std::mutex> g_ctx(c->run_mu);                       // the context's own streams
    Bump bump;
    bump.base = base;
    bump.cap = kNoAllocScratchBytes;
    bwg_batch *b = nullptr;
    rc = stage_ext_bump(c, sc, n, arena, arena_len, q_off, q_len, t_off, t_len, h0, &bump, c->na_ev, c->ss.main, &b);
    if (rc) return rc;
    // the batched form, planned from the tuning table as it stands, launched eagerly
    b->tu.tune_new = false;
    b->tu.use_graphs = false;
    b->tu.form = BWG_FORM_BATCHED;
    rc = enqueue_batch(c, b, BWG_FORM_BATCHED, c->ss);
    if (!rc) rc = enqueue_results(c, b, out, c->ss.main);
    const int rs = bwg_map_hip(hipStreamSynchronize(c->ss.main));
    if (!rc) rc = rs;
    if (!rc) record_run(c, b);
    destroy_batch(c, b);
    return rc;
}

int bwg_extend_side_batch(bwg_ctx *c, const bwg_ext_params *p, int max_band_try, size_t n, const uint8_t *arena,
                          size_t arena_len, const uint64_t *q_off, const uint32_t *q_len, const uint64_t *t_off,
                          const uint32_t *t_len, const int32_t *h0, const int32_t *prev, bwg_ext_result *out,
                          int32_t *w_out) {
    if (!c || !p || max_band_try < 1) return BWG_EINVAL;
    if (n == 0) return BWG_OK;
    if (!out || !w_out || !prev) return BWG_EINVAL;
    const Clock::time_point t0 = Clock::now();
    if (max_band_try > 8 || p->w < 0) return BWG_EINVAL;
    if (p->w > (INT32_MAX >> (max_band_try - 1))) return BWG_E2BIG;
    int rc = check_ext(p, n, arena, arena_len, q_off, q_len, t_off, t_len, h0, host_parts(c->tuning_snapshot(), n));
    if (rc) return rc;
    const ExtScoring sc = ext_scoring(*p);
    return run_on_lane(
        c, t0, Clock::now(),
        [&](hipStream_t st, bwg_batch **b) {
            return stage_side(c, sc, max_band_try, n, arena, arena_len, q_off, q_len, t_off, t_len, h0, prev, st, b,
                              true);
        },
        [&](bwg_batch *b, hipStream_t st) {
            int r = enqueue_results(c, b, out, st);
            return r ? r : fetch_side_bands(c, b, w_out, st);
        });
}

int bwg_extend(bwg_ctx *c, const bwg_ext_params *p, const uint8_t *query, uint32_t qlen,
               const uint8_t *target, uint32_t tlen, int32_t h0, bwg_ext_result *out) {
    if (!c || !p || !out || (!query && qlen) || (!target && tlen)) return BWG_EINVAL;
    std::vector<uint8_t> arena((size_t)qlen + tlen);
    if (qlen) std::memcpy(arena.data(), query, qlen);
    if (tlen) std::memcpy(arena.data() + qlen, target, tlen);
    uint64_t qo = 0, to = qlen;
    return bwg_extend_batch(c, p, 1, arena.data(), arena.size(), &qo, &qlen, &to, &tlen, &h0, out);
}

int bwg_ksw_align2_batch(bwg_ctx *c, const bwg_ksw_params *p, size_t n, const uint8_t *arena,
                         size_t arena_len, const uint64_t *q_off, const uint32_t *q_len,
                         const uint64_t *t_off, const uint32_t *t_len, bwg_ksw_result *out) {
    if (!c || !p) return BWG_EINVAL;
    if (n == 0) return BWG_OK;
    if (!out) return BWG_EINVAL;
    const Clock::time_point t0 = Clock::now();
    ActiveCall active(c);
    if (coalesces(c, active, n)) {
        bwg_chunk ch{};
        ch.kind = BWG_CHUNK_RESCUE;
        ch.ksw = p;
        ch.n = n;
        ch.arena = arena;
        ch.arena_len = arena_len;
        ch.q_off = q_off;
        ch.q_len = q_len;
        ch.t_off = t_off;
        ch.t_len = t_len;
        ch.out = out;
        return run_coalesced(c, ch);
    }
    int rc = check_ksw(p, n, arena, arena_len, q_off, q_len, t_off, t_len, host_parts(c->tuning_snapshot(), n));
    if (rc) return rc;
    const KswScoring sc = ksw_scoring(*p);
    return run_on_lane(
        c, t0, Clock::now(),
        [&](hipStream_t st, bwg_batch **b) {
            return stage_ksw(c, sc, n, arena, arena_len, q_off, q_len, t_off, t_len, nullptr, st, b, true);
        },
        [&](bwg_batch *b, hipStream_t st) { return enqueue_results(c, b, out, st); });
}

int bwg_global_batch(bwg_ctx *c, const bwg_global_params *p, size_t n, const uint8_t *arena, size_t arena_len,
                     const uint64_t *q_off, const uint32_t *q_len, const uint64_t *t_off, const uint32_t *t_len,
                     const int32_t *w, int32_t *score, uint32_t *n_cigar, uint32_t *cigar, size_t cigar_cap) {
    if (!c || !p) return BWG_EINVAL;
    if (n == 0) return BWG_OK;
    if (!score || !n_cigar) return BWG_EINVAL;
    const Clock::time_point t0 = Clock::now();
    int rc = check_global(p, n, arena, arena_len, q_off, q_len, t_off, t_len, w);
    if (rc) return rc;
    const GlobalScoring sc = global_scoring(*p);
    return run_on_lane(
        c, t0, Clock::now(),
        [&](hipStream_t st, bwg_batch **b) {
            return stage_global(c, sc, n, arena, arena_len, q_off, q_len, t_off, t_len, w, st, b);
        },
        [&](bwg_batch *b, hipStream_t st) {
            int r = enqueue_results(c, b, score, st);
            if (!r) r = fetch_cigar_lengths(c, b, n_cigar, st);
            // the run counts are filled even when the runs do not fit cigar_cap (BWG_E2BIG)
            if (!r) r = (!cigar && cigar_cap) ? BWG_EINVAL : fetch_cigars(c, b, cigar, cigar_cap, st);
            return r;
        });
}

int bwg_set_reference(bwg_ctx *c, const uint8_t *fwd, int64_t l_pac, const bwg_contig *contigs,
                      int32_t n_contigs) {
    if (!c || l_pac <= 0 || !fwd || !contigs || n_contigs < 1) return BWG_EINVAL;
    for (int64_t k = 0; k < l_pac; ++k)
        if (fwd[k] > 4) return BWG_EINVAL;
    for (int32_t r = 0; r < n_contigs; ++r) {
        if (contigs[r].offset < 0 || contigs[r].len < 0 || contigs[r].offset + contigs[r].len > l_pac)
            return BWG_EINVAL;
        if (r && contigs[r].offset < contigs[r - 1].offset) return BWG_EINVAL;
    }
    if (hipSetDevice(c->device) != hipSuccess) return BWG_EIO;
    std::lock_guard<std::mutex> g(c->run_mu);
    if (c->ref_users.load()) return BWG_EAGAIN;  // staged rescue batches still read the old one
    if (c->d_ref2) (void)hipFree(c->d_ref2);
    if (c->d_contig_off) (void)hipFree(c->d_contig_off);
    if (c->d_contig_len) (void)hipFree(c->d_contig_len);
    c->d_ref2 = nullptr;
    c->d_contig_off = nullptr;
    c->d_contig_len = nullptr;
    c->l_pac = 0;
    c->n_contigs = 0;
    std::vector<int64_t> off(n_contigs);
    std::vector<int32_t> len(n_contigs);
    for (int32_t r = 0; r < n_contigs; ++r) {
        off[r] = contigs[r].offset;
        len[r] = contigs[r].len;
    }
    hipError_t e = hipMalloc(&c->d_ref2, (size_t)(2 * l_pac) + 64);
    if (e == hipSuccess) e = hipMalloc(&c->d_contig_off, (size_t)n_contigs * 8);
    if (e == hipSuccess) e = hipMalloc(&c->d_contig_len, (size_t)n_contigs * 4);
    if (e == hipSuccess) e = hipMemcpyAsync(c->d_ref2, fwd, (size_t)l_pac, hipMemcpyHostToDevice, c->ss.main);
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->d_contig_off, off.data(), (size_t)n_contigs * 8, hipMemcpyHostToDevice, c->ss.main);
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->d_contig_len, len.data(), (size_t)n_contigs * 4, hipMemcpyHostToDevice, c->ss.main);
    if (e == hipSuccess) e = launch_reverse_half(c->ss.main, c->d_ref2, l_pac);
    if (e == hipSuccess) e = hipStreamSynchronize(c->ss.main);
    if (e != hipSuccess) {
        if (c->d_ref2) (void)hipFree(c->d_ref2);
        if (c->d_contig_off) (void)hipFree(c->d_contig_off);
        if (c->d_contig_len) (void)hipFree(c->d_contig_len);
        c->d_ref2 = nullptr;
        c->d_contig_off = nullptr;
        c->d_contig_len = nullptr;
        return bwg_map_hip(e);
    }
    c->l_pac = l_pac;
    c->n_contigs = n_contigs;
    c->launches += 1;
    return BWG_OK;
}

int bwg_matesw_batch(bwg_ctx *c, const bwg_ksw_params *p, int32_t min_seed_len, const bwg_pe_orient pes[4], size_t n,
                     const bwg_mate_anchor *anchors, const int64_t *ma_rb, size_t n_ma, const uint8_t *arena,
                     size_t arena_len, bwg_rescue_window *win, bwg_ksw_result *out) {
    if (!c || !p || !pes) return BWG_EINVAL;
    if (n == 0) return BWG_OK;
    if (!win || !out || !anchors || (n_ma && !ma_rb)) return BWG_EINVAL;
    const Clock::time_point t0 = Clock::now();
    MateRescueParams mp;
    int rc = check_matesw(c, p, min_seed_len, pes, n, anchors, n_ma, arena, arena_len, &mp);
    if (rc) return rc;
    const KswScoring sc = ksw_scoring(*p);
    return run_on_lane(
        c, t0, Clock::now(),
        [&](hipStream_t st, bwg_batch **b) {
            return stage_matesw(c, sc, mp, n, anchors, ma_rb, n_ma, arena, arena_len, st, b);
        },
        [&](bwg_batch *b, hipStream_t st) {
            int r = enqueue_results(c, b, out, st);
            return r ? r : fetch_windows(c, b, win, st);
        });
}

}  // extern "C"
