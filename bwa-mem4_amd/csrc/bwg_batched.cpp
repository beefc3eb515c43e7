// Batched form: length binning, per-bin kernel choice, persistent work-queue launches, concurrent
// bins, hipGraph replay.
//
// Extension jobs are binned by the work shape that decides the best mapping: short queries run one
// job per lane (ext_thread_kernel), longer ones one job per 16-lane group with K window columns per
// lane (ext_group_kernel<K>), K the smallest instantiated odd value covering the job's widest
// possible window min(qlen, 2w'+1).  Rescue jobs are binned by the padded query columns per lane of
// ksw_group_kernel<K>.  Within a bin, jobs are ordered by descending nominal work so the
// persistent kernels finish with their shortest jobs (longest-processing-time-first).  Jobs outside
// every tuned bin run in the plain kernel.  Each bin's order array and queue head live on the
// device; its kernel pulls from the queue until it is empty.
//
// The bins of one stage have no dependencies on each other, so they are issued on the stream set's
// streams concurrently: a bin's persistent workgroups take over CUs as the previous bin's drain,
// which hides each kernel's tail.  From the second run of a batch on, the whole launch sequence
// (queue reset, every stage's kernels and the fork/join dependencies) is replayed from a captured
// hipGraph.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "bwg_batched_kernels.h"
#include "bwg_runtime.h"

namespace bwg {
namespace {

enum ItemKind { kExtGroup, kExtThread, kExtPlain, kKswFwd, kKswTracker, kKswRev, kKswPlain };

struct Item {
    int kind = 0;
    int param = 0;     // K, or the thread kernel's column capacity
    int qcap = 0;      // extension group: LDS sizing
    int stage = 0;     // items of one stage run concurrently; stages run in order
    uint32_t offset = 0, count = 0, head = 0;
    int grid = 1;
    double work = 0;
};

struct Plan {
    std::vector<Item> items;
    int n_heads = 0;
    uint32_t *d_order = nullptr;
    uint32_t *d_heads = nullptr;
    void *d_jobs = nullptr;  // rescue: descriptors carrying the row-maximum slots
    int32_t *d_rowmax = nullptr;
    void *d_plain_jobs = nullptr;
    void *d_plain_scratch = nullptr;
    uint32_t n_fast = 0;
    // host copies kept alive until the uploads complete
    std::vector<uint32_t> order;
    std::vector<KswJob> ksw_jobs;
    std::vector<ExtJob> ext_plain_jobs;
    std::vector<KswJob> ksw_plain_jobs;
};

int odd_k_for(int width) {
    static const int ks[] = {3, 5, 7, 9, 11, 13, 15, 17, 21, 25, 33};
    for (int k : ks)
        if (16 * k >= width) return k;
    return -1;
}

int qcap_class(int q) { return std::max(32, (q + 31) / 32 * 32); }

int ksw_k_for(int cols) {
    static const int ks[] = {2, 4, 6, 8, 10, 12, 14, 16};
    for (int k : ks)
        if (16 * k >= cols) return k;
    return -1;
}

// Widest padded query either rescue pass of a job can need: the forward pass pads qlen by bwa's
// lane rule, the reverse pass pads qe+1 <= qlen by the same rule, and a shorter prefix can get the
// 16-lane padding where the whole query gets the 8-lane one.
int widest_pad(int qlen, int max_sc) {
    int w = ksw_padded(qlen, max_sc);
    const int x16 = std::min(qlen, 249 / std::max(max_sc, 1));
    if (x16 > 0) w = std::max(w, (x16 + 15) / 16 * 16);
    return w;
}

// Monotone 16-bit key of a job's nominal work: exponent and the 11 bits below the leading one.
inline uint32_t work_key(uint64_t w) {
    if (w == 0) return 0;
    if (w >= (1ull << 31)) w = (1ull << 31) - 1;
    const uint32_t e = 31 - (uint32_t)__builtin_clz((uint32_t)w);
    const uint32_t m = (uint32_t)((w << (31 - e)) >> 20) & 0x7ff;
    return e * 2048 + m + 1;
}

// Group jobs into bins by key, order the bins by stage then descending work and each bin's jobs
// by descending work; returns the order array, bins get their offsets.  Two stable counting-sort
// passes (work key descending, then bin), O(n) for any batch size.
template <class KeyFn, class WorkFn>
std::vector<uint32_t> bin_jobs(size_t n, std::vector<Item> &bins, KeyFn key_of, WorkFn work_of) {
    std::vector<int> key(n);
    std::vector<uint16_t> wk(n);
    for (size_t k = 0; k < n; ++k) {
        key[k] = key_of(k);
        const uint64_t w = work_of(k);
        wk[k] = (uint16_t)(65535 - std::min<uint32_t>(work_key(w), 65535));
    }
    for (size_t k = 0; k < n; ++k) {
        bins[key[k]].count++;
        bins[key[k]].work += (double)work_of(k);
    }
    std::vector<int> perm(bins.size());
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) {
        if (bins[a].stage != bins[b].stage) return bins[a].stage < bins[b].stage;
        return bins[a].work > bins[b].work;
    });
    std::vector<Item> sorted;
    std::vector<int> remap(bins.size());
    for (size_t r = 0; r < perm.size(); ++r) {
        remap[perm[r]] = (int)r;
        sorted.push_back(bins[perm[r]]);
    }
    uint32_t off = 0;
    for (auto &b : sorted) {
        b.offset = off;
        off += b.count;
    }
    // pass 1: by descending work
    std::vector<uint32_t> hist(65537, 0), tmp(n);
    for (size_t k = 0; k < n; ++k) hist[(size_t)wk[k] + 1]++;
    for (size_t v = 1; v < hist.size(); ++v) hist[v] += hist[v - 1];
    for (size_t k = 0; k < n; ++k) tmp[hist[wk[k]]++] = (uint32_t)k;
    // pass 2: stable by bin
    std::vector<uint32_t> order(n), cur(sorted.size());
    for (size_t r = 0; r < sorted.size(); ++r) cur[r] = sorted[r].offset;
    for (size_t r = 0; r < n; ++r) {
        const uint32_t k = tmp[r];
        order[cur[remap[key[k]]]++] = k;
    }
    bins.swap(sorted);
    return order;
}

int plan_ext(bwg_ctx *c, bwg_batch *b, Plan *p, hipStream_t st) {
    const size_t n = b->n;
    const ExtScoring &sc = b->esc;
    const Tuning &tu = c->tuning;
    std::vector<Item> bins;
    auto find = [&](int kind, int param) {
        for (size_t k = 0; k < bins.size(); ++k)
            if (bins[k].kind == kind && bins[k].param == param) return (int)k;
        Item it;
        it.kind = kind;
        it.param = param;
        bins.push_back(it);
        return (int)bins.size() - 1;
    };
    auto key_of = [&](size_t k) {
        const ExtJob &j = b->hext[k];
        const int q = (int)j.qlen;
        const int64_t bound = (int64_t)j.h0 + (int64_t)q * std::max(sc.max_sc, 0);
        if (bound >= (1 << 18) || q >= 8192) return find(kExtPlain, 0);
        if (q <= tu.ext_thread_max_q16) return find(kExtThread, 16);
        if (q <= tu.ext_thread_max_q32) return find(kExtThread, 32);
        const int w = clamp_band(sc.w, q, sc);
        const int K = odd_k_for((int)std::min<int64_t>(q, 2 * (int64_t)w + 1));
        const int qc = qcap_class(q);
        if (K < 0 || qc > tu.ext_group_max_qcap) return find(kExtPlain, 0);
        const int bi = find(kExtGroup, K);
        bins[bi].qcap = std::max(bins[bi].qcap, qc);
        return bi;
    };
    auto work_of = [&](size_t k) { return (uint64_t)b->hext[k].qlen * (b->hext[k].tlen + 1); };
    p->order = bin_jobs(n, bins, key_of, work_of);
    for (size_t r = 0; r < bins.size(); ++r) {
        Item &it = bins[r];
        it.head = (uint32_t)r;
        if (it.kind == kExtGroup) {
            const int per_cu = ext_group_blocks_per_cu(it.param, it.qcap);
            it.grid = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)per_cu * c->num_cus * tu.ext_group_oversub,
                                                                  (it.count + 15) / 16));
        } else if (it.kind == kExtThread) {
            const int per_cu = ext_thread_blocks_per_cu(it.param);
            it.grid = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)per_cu * c->num_cus, (it.count + 63) / 64));
        } else {
            // plain fallback: compact descriptors with per-job scratch offsets
            size_t total = 0;
            for (uint32_t r2 = 0; r2 < it.count; ++r2) {
                ExtJob j = b->hext[p->order[it.offset + r2]];
                j.aux = (uint32_t)total;
                total += (size_t)j.qlen + 1;
                p->ext_plain_jobs.push_back(j);
            }
            if (total >= (1ull << 32)) return BWG_E2BIG;
            p->d_plain_jobs = c->pool.alloc(std::max<size_t>(1, it.count) * sizeof(ExtJob));
            p->d_plain_scratch = c->pool.alloc(std::max<size_t>(1, total) * sizeof(int2));
            if (!p->d_plain_jobs || !p->d_plain_scratch) return BWG_ENOMEM;
        }
    }
    p->items = bins;
    p->n_heads = std::max<int>(1, (int)bins.size());
    p->d_order = (uint32_t *)c->pool.alloc(std::max<size_t>(1, n) * 4);
    p->d_heads = (uint32_t *)c->pool.alloc((size_t)p->n_heads * 4);
    if (!p->d_order || !p->d_heads) return BWG_ENOMEM;
    hipError_t e = hipMemcpyAsync(p->d_order, p->order.data(), n * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && !p->ext_plain_jobs.empty())
        e = hipMemcpyAsync(p->d_plain_jobs, p->ext_plain_jobs.data(), p->ext_plain_jobs.size() * sizeof(ExtJob),
                           hipMemcpyHostToDevice, st);
    return bwg_map_hip(e);
}

int plan_ksw(bwg_ctx *c, bwg_batch *b, Plan *p, hipStream_t st) {
    const size_t n = b->n;
    const KswScoring &sc = b->ksc;
    const Tuning &tu = c->tuning;
    std::vector<Item> bins;
    auto find = [&](int kind, int param) {
        for (size_t k = 0; k < bins.size(); ++k)
            if (bins[k].kind == kind && bins[k].param == param) return (int)k;
        Item it;
        it.kind = kind;
        it.param = param;
        it.stage = kind == kKswFwd ? 0 : 1;  // fast bins first in the order array
        bins.push_back(it);
        return (int)bins.size() - 1;
    };
    auto key_of = [&](size_t k) {
        const int K = sc.o_ins >= 0 ? ksw_k_for(widest_pad((int)b->hksw[k].qlen, sc.max_sc)) : -1;
        return K < 0 ? find(kKswPlain, 0) : find(kKswFwd, K);
    };
    auto work_of = [&](size_t k) { return (uint64_t)b->hksw[k].tlen; };
    p->order = bin_jobs(n, bins, key_of, work_of);
    // Row-maximum slots for the fast-path jobs: tiles of 64 consecutive jobs of the order array,
    // row-major within a tile ([row][64]), so the tracker kernel's 64 lanes read consecutive words.
    // A job's slot for row i is aux0 + 64*i.
    p->ksw_jobs = b->hksw;
    for (auto &it : bins)
        if (it.kind == kKswFwd) p->n_fast += it.count;
    uint64_t rows = 0;
    for (uint32_t t0 = 0; t0 < p->n_fast; t0 += 64) {
        uint32_t tmax = 0;
        const uint32_t t1 = std::min<uint32_t>(p->n_fast, t0 + 64);
        for (uint32_t r = t0; r < t1; ++r) tmax = std::max(tmax, p->ksw_jobs[p->order[r]].tlen);
        for (uint32_t r = t0; r < t1; ++r) p->ksw_jobs[p->order[r]].aux0 = (uint32_t)(rows + (r - t0));
        rows += 64ull * tmax;
    }
    if (rows >= (1ull << 32)) return BWG_E2BIG;
    std::vector<Item> items;
    int heads = 0;
    for (auto &it : bins) {
        if (it.kind == kKswFwd) {
            Item f = it;
            f.stage = 0;
            f.head = (uint32_t)heads++;
            const int gf = ksw_group_blocks_per_cu(it.param, false);
            f.grid = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)gf * c->num_cus * tu.ksw_group_oversub,
                                                                 (it.count + 15) / 16));
            items.push_back(f);
            Item r = it;
            r.kind = kKswRev;
            r.stage = 2;
            r.head = (uint32_t)heads++;
            const int gr = ksw_group_blocks_per_cu(it.param, true);
            r.grid = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)gr * c->num_cus * tu.ksw_group_oversub,
                                                                 (it.count + 15) / 16));
            items.push_back(r);
        } else {
            Item pl = it;
            pl.stage = 0;
            size_t total = 0;
            for (uint32_t r2 = 0; r2 < it.count; ++r2) {
                KswJob j = b->hksw[p->order[it.offset + r2]];
                const size_t bytes = ((size_t)ksw_padded((int32_t)j.qlen, sc.max_sc) + 16) * 8 + (size_t)j.tlen * 4;
                j.aux0 = (uint32_t)(total / 16);
                total += (bytes + 15) & ~(size_t)15;
                p->ksw_plain_jobs.push_back(j);
            }
            if (total / 16 >= (1ull << 32)) return BWG_E2BIG;
            p->d_plain_jobs = c->pool.alloc(std::max<size_t>(1, it.count) * sizeof(KswJob));
            p->d_plain_scratch = c->pool.alloc(std::max<size_t>(16, total));
            if (!p->d_plain_jobs || !p->d_plain_scratch) return BWG_ENOMEM;
            items.push_back(pl);
        }
    }
    if (p->n_fast) {
        Item t;
        t.kind = kKswTracker;
        t.stage = 1;
        t.count = p->n_fast;
        items.push_back(t);
    }
    std::stable_sort(items.begin(), items.end(), [](const Item &a, const Item &b2) { return a.stage < b2.stage; });
    p->items = items;
    p->n_heads = std::max(1, heads);
    p->d_order = (uint32_t *)c->pool.alloc(std::max<size_t>(1, n) * 4);
    p->d_heads = (uint32_t *)c->pool.alloc((size_t)p->n_heads * 4);
    p->d_jobs = c->pool.alloc(std::max<size_t>(1, n) * sizeof(KswJob));
    p->d_rowmax = (int32_t *)c->pool.alloc(std::max<size_t>(1, rows) * 4);
    if (!p->d_order || !p->d_heads || !p->d_jobs || !p->d_rowmax) return BWG_ENOMEM;
    hipError_t e = hipMemcpyAsync(p->d_order, p->order.data(), n * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(p->d_jobs, p->ksw_jobs.data(), n * sizeof(KswJob), hipMemcpyHostToDevice, st);
    if (e == hipSuccess && !p->ksw_plain_jobs.empty())
        e = hipMemcpyAsync(p->d_plain_jobs, p->ksw_plain_jobs.data(), p->ksw_plain_jobs.size() * sizeof(KswJob),
                           hipMemcpyHostToDevice, st);
    return bwg_map_hip(e);
}

hipError_t launch_item(bwg_batch *b, Plan *p, const Item &it, hipStream_t st) {
    switch (it.kind) {
    case kExtGroup:
        return launch_ext_group(st, it.param, b->d_arena, b->d_ref, (const ExtJob *)b->d_jobs, p->d_order + it.offset,
                                it.count, p->d_heads + it.head, b->esc, it.qcap, (bwg_ext_result *)b->d_out, it.grid);
    case kExtThread:
        return launch_ext_thread(st, it.param, b->d_arena, b->d_ref, (const ExtJob *)b->d_jobs, p->d_order + it.offset,
                                 it.count, p->d_heads + it.head, b->esc, (bwg_ext_result *)b->d_out, it.grid);
    case kExtPlain:
        return launch_ext_plain_scatter(st, b->d_arena, b->d_ref, (const ExtJob *)p->d_plain_jobs,
                                        p->d_order + it.offset, it.count, b->esc, (int2 *)p->d_plain_scratch,
                                        (bwg_ext_result *)b->d_out);
    case kKswFwd:
    case kKswRev:
        return launch_ksw_group(st, it.param, it.kind == kKswRev, b->d_arena, b->d_ref, (const KswJob *)p->d_jobs,
                                p->d_order + it.offset, it.count, p->d_heads + it.head, b->ksc,
                                (bwg_ksw_result *)b->d_out, p->d_rowmax, it.grid);
    case kKswTracker:
        return launch_ksw_tracker(st, (const KswJob *)p->d_jobs, p->d_order, it.count, b->ksc, p->d_rowmax,
                                  (bwg_ksw_result *)b->d_out);
    case kKswPlain:
        return launch_ksw_plain_scatter(st, b->d_arena, b->d_ref, (const KswJob *)p->d_plain_jobs,
                                        p->d_order + it.offset, it.count, b->ksc, (uint8_t *)p->d_plain_scratch,
                                        (bwg_ksw_result *)b->d_out);
    }
    return hipErrorInvalidValue;
}

// The launch sequence of a planned batch on a stream set: queue reset, then each stage's items,
// concurrently over the streams when enabled.
hipError_t emit(bwg_ctx *c, bwg_batch *b, Plan *p, StreamSet &ss, int64_t *launches) {
    hipError_t e = launch_zero_u32(ss.main, p->d_heads, (uint32_t)p->n_heads);
    int64_t nl = 1;
    size_t i = 0;
    while (i < p->items.size() && e == hipSuccess) {
        size_t j = i;
        while (j < p->items.size() && p->items[j].stage == p->items[i].stage) ++j;
        const int nitems = (int)(j - i);
        const bool par = c->tuning.concurrent_bins && nitems > 1;
        if (par) {
            e = hipEventRecord(ss.fork, ss.main);
            for (int k = 0; k < kAuxStreams && e == hipSuccess; ++k) e = hipStreamWaitEvent(ss.aux[k], ss.fork, 0);
        }
        for (size_t r = i; r < j && e == hipSuccess; ++r) {
            const int slot = (int)((r - i) % (kAuxStreams + 1));
            hipStream_t st = (!par || slot == 0) ? ss.main : ss.aux[slot - 1];
            e = launch_item(b, p, p->items[r], st);
            ++nl;
        }
        if (par) {
            for (int k = 0; k < kAuxStreams && e == hipSuccess; ++k) {
                e = hipEventRecord(ss.join[k], ss.aux[k]);
                if (e == hipSuccess) e = hipStreamWaitEvent(ss.main, ss.join[k], 0);
            }
        }
        i = j;
    }
    *launches = nl;
    return e;
}

}  // namespace

int enqueue_batched(bwg_ctx *c, bwg_batch *b, StreamSet &ss, int64_t *launches) {
    *launches = 0;
    if (!b->plan) {
        Plan *p = new Plan;
        b->plan = p;
        const int rc = b->kind == 0 ? plan_ext(c, b, p, ss.main) : plan_ksw(c, b, p, ss.main);
        if (rc) return rc;
    }
    Plan *p = (Plan *)b->plan;
    hipError_t e = hipSuccess;
    // Capture the launch sequence once the batch has run eagerly; the eager run is also the
    // reference every later replay is compared with by the tests.
    if (c->tuning.use_graphs && !b->graph && b->batched_runs >= 1) {
        hipGraph_t g = nullptr;
        e = hipStreamBeginCapture(ss.main, hipStreamCaptureModeThreadLocal);
        int64_t nl = 0;
        const hipError_t e2 = e == hipSuccess ? emit(c, b, p, ss, &nl) : e;
        if (e == hipSuccess) e = hipStreamEndCapture(ss.main, &g);
        if (e == hipSuccess && e2 != hipSuccess) e = e2;
        if (e == hipSuccess) e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
        if (g) (void)hipGraphDestroy(g);
        if (e != hipSuccess) return bwg_map_hip(e);
    }
    e = hipEventRecord(b->ev_start, ss.main);
    if (e == hipSuccess) {
        if (b->graph) {
            e = hipGraphLaunch(b->graph, ss.main);
            *launches = (int64_t)p->items.size() + 1;
        } else {
            e = emit(c, b, p, ss, launches);
        }
    }
    if (e == hipSuccess) e = hipEventRecord(b->ev_end, ss.main);
    b->batched_runs++;
    return bwg_map_hip(e);
}

void release_batched(bwg_ctx *c, bwg_batch *b) {
    if (b->graph) {
        (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    if (!b->plan) return;
    Plan *p = (Plan *)b->plan;
    c->pool.free(p->d_order);
    c->pool.free(p->d_heads);
    c->pool.free(p->d_jobs);
    c->pool.free(p->d_rowmax);
    c->pool.free(p->d_plain_jobs);
    c->pool.free(p->d_plain_scratch);
    delete p;
    b->plan = nullptr;
}

}  // namespace bwg
