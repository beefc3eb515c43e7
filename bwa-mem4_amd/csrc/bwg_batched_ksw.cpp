// Batched form of mate rescue: jobs binned by the number of padded query columns each lane of the
// 16-lane group kernel owns (K), each bin ordered by descending target length; forward pass,
// tracker replay, reverse start-recovery pass.  Jobs outside the tuned bins (padded query over
// 256 columns, or a negative insertion-open penalty, for which the lazy-F reduction does not hold)
// run in the plain kernel.
#include <algorithm>
#include <numeric>
#include <vector>

#include "bwg_batched.h"
#include "bwg_batched_kernels.h"
#include "bwg_tuning.h"

namespace bwg {
namespace {

struct KswBin {
    int K = 0;  // 0: plain fallback
    uint32_t offset = 0, count = 0;
};

struct KswPlan {
    std::vector<KswBin> bins;
    uint32_t *d_order = nullptr;   // all jobs, bin by bin
    uint32_t *d_heads = nullptr;   // two queue heads (forward, reverse) per bin
    KswJob *d_jobs = nullptr;      // descriptors with the row-maximum offsets
    int32_t *d_rowmax = nullptr;
    uint32_t n_fast = 0;
    // plain fallback
    KswJob *d_plain_jobs = nullptr;
    uint8_t *d_plain_scratch = nullptr;
};

int ksw_k_for(int cols) {
    static const int ks[] = {2, 4, 6, 8, 10, 12, 14, 16};
    for (int k : ks)
        if (16 * k >= cols) return k;
    return -1;
}

// Widest padded query either pass of a job can need: the forward pass pads qlen by bwa's lane
// rule, the reverse pass pads qe+1 <= qlen by the same rule, and a shorter prefix can get the
// 16-lane padding where the whole query gets the 8-lane one.
int widest_pad(int qlen, int max_sc) {
    int w = ksw_padded(qlen, max_sc);
    int x16 = std::min(qlen, 249 / std::max(max_sc, 1));
    if (x16 > 0) w = std::max(w, (x16 + 15) / 16 * 16);
    return w;
}

KswPlan *build_plan(bwg_ctx *c, bwg_batch *b, int *rc) {
    const size_t n = b->n;
    const KswScoring &sc = b->ksc;
    KswPlan *p = new KswPlan;
    std::vector<int> kof(n);
    std::vector<KswBin> bins;
    std::vector<int> bin_of_k(17, -1);
    int plain_bin = -1;
    for (size_t k = 0; k < n; ++k) {
        int K = sc.o_ins >= 0 ? ksw_k_for(widest_pad((int)b->q_len[k], sc.max_sc)) : -1;
        int bi;
        if (K < 0) {
            if (plain_bin < 0) {
                plain_bin = (int)bins.size();
                bins.push_back(KswBin());
            }
            bi = plain_bin;
        } else {
            if (bin_of_k[K] < 0) {
                bin_of_k[K] = (int)bins.size();
                KswBin nb;
                nb.K = K;
                bins.push_back(nb);
            }
            bi = bin_of_k[K];
        }
        kof[k] = bi;
        bins[bi].count++;
    }
    // fast bins first (widest first), the plain fallback last
    std::vector<int> perm(bins.size());
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int c2) {
        if ((bins[a].K == 0) != (bins[c2].K == 0)) return bins[a].K != 0;
        return bins[a].K > bins[c2].K;
    });
    std::vector<int> remap(bins.size());
    for (size_t r = 0; r < perm.size(); ++r) {
        remap[perm[r]] = (int)r;
        p->bins.push_back(bins[perm[r]]);
    }
    uint32_t off = 0;
    for (auto &bn : p->bins) {
        bn.offset = off;
        off += bn.count;
    }
    std::vector<uint32_t> order(n), cur(p->bins.size());
    for (size_t r = 0; r < p->bins.size(); ++r) cur[r] = p->bins[r].offset;
    for (size_t k = 0; k < n; ++k) order[cur[remap[kof[k]]]++] = (uint32_t)k;
    for (auto &bn : p->bins)
        std::sort(order.begin() + bn.offset, order.begin() + bn.offset + bn.count, [&](uint32_t a, uint32_t c2) {
            return b->t_len[a] != b->t_len[c2] ? b->t_len[a] > b->t_len[c2] : a < c2;
        });
    std::vector<KswJob> all(n);
    hipError_t e = hipMemcpy(all.data(), b->d_jobs, n * sizeof(KswJob), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        *rc = bwg_map_hip(e);
        return p;
    }
    // Row-maximum slots for the fast-path jobs: tiles of 64 consecutive jobs of the order array,
    // row-major within a tile ([row][64]), so the tracker kernel's 64 lanes read consecutive words.
    // A job's slot for row i is aux0 + 64*i.
    uint64_t rows = 0;
    std::vector<KswJob> plain;
    for (auto &bn : p->bins) {
        if (bn.K) p->n_fast += bn.count;
        else
            for (uint32_t r = 0; r < bn.count; ++r) plain.push_back(all[order[bn.offset + r]]);
    }
    for (uint32_t t0 = 0; t0 < p->n_fast; t0 += 64) {
        uint32_t tmax = 0;
        const uint32_t t1 = std::min<uint32_t>(p->n_fast, t0 + 64);
        for (uint32_t r = t0; r < t1; ++r) tmax = std::max(tmax, all[order[r]].tlen);
        for (uint32_t r = t0; r < t1; ++r) all[order[r]].aux0 = (uint32_t)(rows + (r - t0));
        rows += 64ull * tmax;
    }
    if (rows >= (1ull << 32)) {
        *rc = BWG_E2BIG;
        return p;
    }
    size_t total = 0;
    for (auto &pj : plain) {
        size_t bytes = ((size_t)ksw_padded((int32_t)pj.qlen, sc.max_sc) + 16) * 8 + (size_t)pj.tlen * 4;
        pj.aux0 = (uint32_t)(total / 16);
        total += (bytes + 15) & ~(size_t)15;
    }
    p->d_order = (uint32_t *)c->pool.alloc(std::max<size_t>(1, n) * 4);
    p->d_heads = (uint32_t *)c->pool.alloc(std::max<size_t>(2, 2 * p->bins.size()) * 4);
    p->d_jobs = (KswJob *)c->pool.alloc(std::max<size_t>(1, n) * sizeof(KswJob));
    p->d_rowmax = (int32_t *)c->pool.alloc(std::max<size_t>(1, rows) * 4);
    if (!plain.empty()) {
        p->d_plain_jobs = (KswJob *)c->pool.alloc(plain.size() * sizeof(KswJob));
        p->d_plain_scratch = (uint8_t *)c->pool.alloc(std::max<size_t>(16, total));
    }
    if (!p->d_order || !p->d_heads || !p->d_jobs || !p->d_rowmax ||
        (!plain.empty() && (!p->d_plain_jobs || !p->d_plain_scratch))) {
        *rc = BWG_ENOMEM;
        return p;
    }
    e = hipMemcpy(p->d_order, order.data(), n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->d_jobs, all.data(), n * sizeof(KswJob), hipMemcpyHostToDevice);
    if (e == hipSuccess && !plain.empty())
        e = hipMemcpy(p->d_plain_jobs, plain.data(), plain.size() * sizeof(KswJob), hipMemcpyHostToDevice);
    *rc = bwg_map_hip(e);
    return p;
}

}  // namespace

int run_batched_ksw(bwg_ctx *c, bwg_batch *b, BatchedTiming *t) {
    int rc = BWG_OK;
    if (!b->ksw_plan) {
        b->ksw_plan = build_plan(c, b, &rc);
        if (rc) return rc;
    }
    KswPlan *p = (KswPlan *)b->ksw_plan;
    const Tuning &tu = tuning_for(c);
    bwg_ksw_result *out = (bwg_ksw_result *)b->d_out;
    hipError_t e = hipEventRecord(c->ev[0], c->stream);
    if (e == hipSuccess) e = launch_zero_u32(c->stream, p->d_heads, (uint32_t)(2 * p->bins.size()));
    int64_t launches = 1, main_launches = 0;
    if (e == hipSuccess && p->n_fast) e = hipEventRecord(c->ev[2], c->stream);
    for (size_t r = 0; r < p->bins.size() && e == hipSuccess; ++r) {
        const KswBin &bn = p->bins[r];
        if (!bn.K || !bn.count) continue;
        int grid = std::min<int64_t>((int64_t)ksw_group_blocks_per_cu(bn.K, false) * c->num_cus * tu.ksw_group_oversub,
                                     (bn.count + 15) / 16);
        e = launch_ksw_group(c->stream, bn.K, false, b->d_arena, p->d_jobs, p->d_order + bn.offset, bn.count,
                             p->d_heads + 2 * r, b->ksc, out, p->d_rowmax, std::max(grid, 1));
        ++launches;
        ++main_launches;
    }
    if (e == hipSuccess && p->n_fast) e = hipEventRecord(c->ev[3], c->stream);
    if (e == hipSuccess && p->n_fast) {
        e = launch_ksw_tracker(c->stream, p->d_jobs, p->d_order, p->n_fast, b->ksc, p->d_rowmax, out);
        ++launches;
    }
    for (size_t r = 0; r < p->bins.size() && e == hipSuccess; ++r) {
        const KswBin &bn = p->bins[r];
        if (!bn.K || !bn.count) continue;
        int grid = std::min<int64_t>((int64_t)ksw_group_blocks_per_cu(bn.K, true) * c->num_cus * tu.ksw_group_oversub,
                                     (bn.count + 15) / 16);
        e = launch_ksw_group(c->stream, bn.K, true, b->d_arena, p->d_jobs, p->d_order + bn.offset, bn.count,
                             p->d_heads + 2 * r + 1, b->ksc, out, p->d_rowmax, std::max(grid, 1));
        ++launches;
    }
    for (size_t r = 0; r < p->bins.size() && e == hipSuccess; ++r) {
        const KswBin &bn = p->bins[r];
        if (bn.K || !bn.count) continue;
        e = launch_ksw_plain_scatter(c->stream, b->d_arena, p->d_plain_jobs, p->d_order + bn.offset, bn.count,
                                     b->ksc, p->d_plain_scratch, out);
        ++launches;
    }
    if (e == hipSuccess) e = hipEventRecord(c->ev[1], c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return bwg_map_hip(e);
    float ms = 0.f, mms = 0.f;
    (void)hipEventElapsedTime(&ms, c->ev[0], c->ev[1]);
    if (p->n_fast) (void)hipEventElapsedTime(&mms, c->ev[2], c->ev[3]);
    t->total_ns = (int64_t)(ms * 1e6);
    t->main_ns = (int64_t)(mms * 1e6);
    t->launches = launches;
    t->main_launches = main_launches;
    return BWG_OK;
}

void release_batched_ksw(bwg_ctx *c, bwg_batch *b) {
    if (!b->ksw_plan) return;
    KswPlan *p = (KswPlan *)b->ksw_plan;
    c->pool.free(p->d_order);
    c->pool.free(p->d_heads);
    c->pool.free(p->d_jobs);
    c->pool.free(p->d_rowmax);
    c->pool.free(p->d_plain_jobs);
    c->pool.free(p->d_plain_scratch);
    delete p;
    b->ksw_plan = nullptr;
}

const Tuning &tuning_for(bwg_ctx *) {
    static Tuning t;
    return t;
}

}  // namespace bwg
