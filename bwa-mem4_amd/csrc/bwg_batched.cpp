// Batched form: length binning, per-bin kernel choice, persistent work-queue launches.
//
// Jobs are binned by the work shape that decides the best mapping: short queries run one job per
// lane (ext_thread_kernel), longer ones one job per 16-lane group with K window columns per lane
// (ext_group_kernel<K>), K the smallest instantiated odd value covering the job's widest possible
// window min(qlen, 2w'+1).  Within a bin, jobs are ordered by descending nominal work so the
// persistent kernels finish with their shortest jobs (longest-processing-time-first).  Jobs that
// exceed every tuned bin run in the plain kernel.  Each bin's order array and queue head live on
// the device; the kernels pull from the queue until it is empty.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "bwg_batched.h"
#include "bwg_batched_kernels.h"
#include "bwg_tuning.h"

namespace bwg {

namespace {

enum BinKind { kThread = 0, kGroup = 1, kPlain = 2 };

struct Bin {
    int kind = kGroup;
    int param = 0;  // thread: qcap; group: K
    int qcap = 0;   // group: bin key part (0: one bin per K)
    int qcap_max = 0;  // group: LDS sizing, the longest query of the bin rounded up to 32
    uint32_t offset = 0, count = 0;
    double work = 0;
};

struct ExtPlan {
    std::vector<Bin> bins;
    uint32_t *d_order = nullptr;
    uint32_t *d_heads = nullptr;
    ExtJob *d_plain_jobs = nullptr;  // fallback jobs with scratch offsets
    int2 *d_plain_scratch = nullptr;
    uint32_t n_plain = 0;
};

int odd_k_for(int width) {
    static const int ks[] = {3, 5, 7, 9, 11, 13, 15, 17, 21, 25, 33};
    for (int k : ks)
        if (16 * k >= width) return k;
    return -1;
}

int qcap_class(int q) { return std::max(32, (q + 31) / 32 * 32); }

ExtPlan *build_ext_plan(bwg_ctx *c, bwg_batch *b, int *rc) {
    const size_t n = b->n;
    const ExtScoring &sc = b->esc;
    const Tuning &tu = tuning_for(c);
    // bin key per job
    std::vector<int> key(n);
    std::vector<Bin> bins;
    auto find_bin = [&](int kind, int param, int qcap) {
        for (size_t k = 0; k < bins.size(); ++k)
            if (bins[k].kind == kind && bins[k].param == param && bins[k].qcap == qcap) return (int)k;
        Bin nb;
        nb.kind = kind;
        nb.param = param;
        nb.qcap = qcap;
        bins.push_back(nb);
        return (int)bins.size() - 1;
    };
    for (size_t k = 0; k < n; ++k) {
        const int q = (int)b->q_len[k];
        const int64_t bound = (int64_t)b->h0[k] + (int64_t)q * std::max(sc.max_sc, 0);
        const int w = clamp_band(sc.w, q, sc);
        const int width = std::min<int64_t>(q, 2 * (int64_t)w + 1);
        int bi;
        if (bound >= (1 << 18) || q >= 8192) {
            bi = find_bin(kPlain, 0, 0);
        } else if (q <= tu.ext_thread_max_q16) {
            bi = find_bin(kThread, 16, 0);
        } else if (q <= tu.ext_thread_max_q32) {
            bi = find_bin(kThread, 32, 0);
        } else {
            const int K = odd_k_for(width);
            const int qc = qcap_class(q);
            if (K < 0 || qc > tu.ext_group_max_qcap) bi = find_bin(kPlain, 0, 0);
            else bi = find_bin(kGroup, K, 0);
            if (bi >= 0 && bins[bi].kind == kGroup) bins[bi].qcap_max = std::max(bins[bi].qcap_max, qc);
        }
        key[k] = bi;
        bins[bi].count++;
        bins[bi].work += (double)q * (double)b->t_len[k];
    }
    // order bins: group bins (most work first), then thread bins, then the plain fallback
    std::vector<int> border(bins.size());
    std::iota(border.begin(), border.end(), 0);
    std::stable_sort(border.begin(), border.end(), [&](int a, int c2) {
        if (bins[a].kind != bins[c2].kind) return bins[a].kind == kGroup || (bins[a].kind == kThread && bins[c2].kind == kPlain);
        return bins[a].work > bins[c2].work;
    });
    std::vector<int> remap(bins.size());
    std::vector<Bin> sorted_bins;
    for (size_t r = 0; r < border.size(); ++r) {
        remap[border[r]] = (int)r;
        sorted_bins.push_back(bins[border[r]]);
    }
    uint32_t off = 0;
    for (auto &bn : sorted_bins) {
        bn.offset = off;
        off += bn.count;
    }
    // fill the order array: each bin's jobs by descending nominal work (qlen*tlen)
    std::vector<uint32_t> order(n);
    std::vector<uint32_t> cursor(sorted_bins.size());
    for (size_t r = 0; r < sorted_bins.size(); ++r) cursor[r] = sorted_bins[r].offset;
    for (size_t k = 0; k < n; ++k) order[cursor[remap[key[k]]]++] = (uint32_t)k;
    for (auto &bn : sorted_bins) {
        std::sort(order.begin() + bn.offset, order.begin() + bn.offset + bn.count, [&](uint32_t a, uint32_t c2) {
            uint64_t wa = (uint64_t)b->q_len[a] * (b->t_len[a] + 1), wc = (uint64_t)b->q_len[c2] * (b->t_len[c2] + 1);
            return wa != wc ? wa > wc : a < c2;
        });
    }
    ExtPlan *p = new ExtPlan;
    p->bins = sorted_bins;
    p->d_order = (uint32_t *)c->pool.alloc(std::max<size_t>(1, n) * 4);
    p->d_heads = (uint32_t *)c->pool.alloc(std::max<size_t>(1, p->bins.size()) * 4);
    if (!p->d_order || !p->d_heads) {
        *rc = BWG_ENOMEM;
        return p;
    }
    hipError_t e = hipMemcpy(p->d_order, order.data(), n * 4, hipMemcpyHostToDevice);
    // the plain fallback needs per-job scratch offsets
    for (auto &bn : p->bins) {
        if (bn.kind != kPlain || bn.count == 0 || e != hipSuccess) continue;
        std::vector<ExtJob> pj(bn.count);
        std::vector<ExtJob> all(n);
        e = hipMemcpy(all.data(), b->d_jobs, n * sizeof(ExtJob), hipMemcpyDeviceToHost);
        size_t total = 0;
        for (uint32_t r = 0; r < bn.count; ++r) {
            pj[r] = all[order[bn.offset + r]];
            pj[r].aux = (uint32_t)total;
            total += (size_t)pj[r].qlen + 1;
        }
        // results of the fallback are written to a compact array and scattered back by index
        p->n_plain = bn.count;
        p->d_plain_jobs = (ExtJob *)c->pool.alloc(bn.count * sizeof(ExtJob));
        p->d_plain_scratch = (int2 *)c->pool.alloc(std::max<size_t>(1, total) * sizeof(int2));
        if (!p->d_plain_jobs || !p->d_plain_scratch) {
            *rc = BWG_ENOMEM;
            return p;
        }
        if (e == hipSuccess)
            e = hipMemcpy(p->d_plain_jobs, pj.data(), bn.count * sizeof(ExtJob), hipMemcpyHostToDevice);
    }
    *rc = bwg_map_hip(e);
    return p;
}

}  // namespace

int run_batched(bwg_ctx *c, bwg_batch *b, BatchedTiming *t) {
    if (b->kind != 0) return run_batched_ksw(c, b, t);
    int rc = BWG_OK;
    if (!b->ext_plan) {
        b->ext_plan = build_ext_plan(c, b, &rc);
        if (rc) return rc;
    }
    ExtPlan *p = (ExtPlan *)b->ext_plan;
    const Tuning &tu = tuning_for(c);
    hipError_t e = hipEventRecord(c->ev[0], c->stream);
    if (e == hipSuccess) e = launch_zero_u32(c->stream, p->d_heads, (uint32_t)p->bins.size());
    int64_t launches = 1, main_launches = 0;
    bool main_open = false;
    bwg_ext_result *out = (bwg_ext_result *)b->d_out;
    for (size_t r = 0; r < p->bins.size() && e == hipSuccess; ++r) {
        const Bin &bn = p->bins[r];
        if (bn.count == 0) continue;
        if (bn.kind == kGroup) {
            if (!main_open) {
                e = hipEventRecord(c->ev[2], c->stream);
                main_open = true;
            }
            const int per_cu = ext_group_blocks_per_cu(bn.param, bn.qcap_max);
            const uint32_t groups_needed = bn.count;
            int grid = std::min<int64_t>((int64_t)per_cu * c->num_cus * tu.ext_group_oversub, (groups_needed + 15) / 16);
            grid = std::max(grid, 1);
            if (e == hipSuccess)
                e = launch_ext_group(c->stream, bn.param, b->d_arena, (const ExtJob *)b->d_jobs,
                                     p->d_order + bn.offset, bn.count, p->d_heads + r, b->esc, bn.qcap_max,
                                     out, grid);
            ++launches;
            ++main_launches;
        } else {
            if (main_open) {
                e = hipEventRecord(c->ev[3], c->stream);
                main_open = false;
            }
            if (bn.kind == kThread) {
                const int per_cu = ext_thread_blocks_per_cu(bn.param);
                int grid = std::min<int64_t>((int64_t)per_cu * c->num_cus, (bn.count + 63) / 64);
                grid = std::max(grid, 1);
                if (e == hipSuccess)
                    e = launch_ext_thread(c->stream, bn.param, b->d_arena, (const ExtJob *)b->d_jobs,
                                          p->d_order + bn.offset, bn.count, p->d_heads + r, b->esc, out, grid);
            } else {
                if (e == hipSuccess)
                    e = launch_ext_plain_scatter(c->stream, b->d_arena, p->d_plain_jobs, p->d_order + bn.offset,
                                                 bn.count, b->esc, p->d_plain_scratch, out);
            }
            ++launches;
        }
    }
    if (main_open && e == hipSuccess) e = hipEventRecord(c->ev[3], c->stream);
    if (e == hipSuccess) e = hipEventRecord(c->ev[1], c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return bwg_map_hip(e);
    float ms = 0.f, mms = 0.f;
    (void)hipEventElapsedTime(&ms, c->ev[0], c->ev[1]);
    if (main_launches) (void)hipEventElapsedTime(&mms, c->ev[2], c->ev[3]);
    t->total_ns = (int64_t)(ms * 1e6);
    t->main_ns = (int64_t)(mms * 1e6);
    t->launches = launches;
    t->main_launches = main_launches;
    return BWG_OK;
}

void release_batched(bwg_ctx *c, bwg_batch *b) {
    if (b->ext_plan) {
        ExtPlan *p = (ExtPlan *)b->ext_plan;
        c->pool.free(p->d_order);
        c->pool.free(p->d_heads);
        c->pool.free(p->d_plain_jobs);
        c->pool.free(p->d_plain_scratch);
        delete p;
        b->ext_plan = nullptr;
    }
    release_batched_ksw(c, b);
}

}  // namespace bwg
