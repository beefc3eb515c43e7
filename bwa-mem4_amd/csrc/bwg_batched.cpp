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
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <vector>

#include "bwg_batched_kernels.h"
#include "bwg_runtime.h"
#include "bwg_tuner.h"

namespace bwg {
namespace {

struct Item {
    int kind = 0;
    int param = 0;     // K, or the thread kernel's column capacity
    bool narrow = false;  // 16-bit DP state (extension, packed rescue kernel); packed scores (global)
    int qcap = 0;      // extension group: LDS sizing
    int stage = 0;     // items of one stage run concurrently; stages run in order
    uint32_t offset = 0, count = 0, head = 0;
    int grid = 1;
    int launch_grid = 0;  // grid of the last launch sequence issued (emit may size it below grid)
    double work = 0;
    double est_ns = 0;     // extension DP bin: the autotuner's time per job of each job's class, summed
    bool est_all = false;  // every job of the bin has a tuned class (est_ns is its whole time alone)
    double dp_share = -1;  // extension DP bin after the pre-pass: measured share of its jobs the DP runs
};

// Bin code of an item, the key of the context's measured DP shares.
inline int item_code(const Item &it) { return (it.kind << 8) | ((it.param & 127) << 1) | (it.narrow ? 1 : 0); }

struct Plan {
    ExtScoring esc{};      // extension: the scoring this plan's kernels run with (its band w)
    int retry_round = 0;   // extension: 0, or the band-retry round this plan runs
    bool compact = false;  // group and thread bins drain compacted queues filled by a stage-0 item
    std::vector<Item> items;
    std::vector<hipEvent_t> ev0, ev1;  // per-item timing events (item_timing), created on first use
    int n_heads = 0;
    uint32_t *d_order = nullptr;
    uint32_t *d_heads = nullptr;
    void *d_jobs = nullptr;  // rescue: descriptors carrying the row-maximum slots
    int32_t *d_rowmax = nullptr;
    void *d_plain_jobs = nullptr;
    void *d_plain_scratch = nullptr;
    int32_t *d_plain_slot = nullptr;  // rescue windows: a job's index among the plain descriptors, or -1
    uint32_t n_fast = 0;
    // extension ungapped pre-pass: compacted per-bin queues; d_heads holds the queue heads, then the
    // compacted counts, then the bins' slot offsets and skip flags
    bool ungapped = false;
    int32_t ungapped_xmax = -1;
    uint32_t *d_corder = nullptr;
    uint8_t *d_declined = nullptr;  // the pre-pass verdict per job
    int n_bins = 0;
    std::vector<uint32_t> meta;
    // host copies kept alive until the uploads complete
    HostVec<uint32_t> order;
    std::vector<KswJob> ksw_jobs;
    std::vector<ExtJob> ext_plain_jobs;
    std::vector<KswJob> ksw_plain_jobs;
};

// First word of the slot offsets in the extension plan's metadata (after the queue heads and the
// spaced-out compacted-queue counters).
inline size_t seg_word(const Plan *p) { return (size_t)p->n_bins * (1 + kCountStride); }

// Autotuning: a class is tuned once the context sees a batch with at least kTuneMinJobs of its
// jobs; each candidate runs kTuneReps times on up to kTuneSample of them and keeps its best time.
constexpr size_t kTuneMinJobs = 4096;
constexpr size_t kTuneSample = 65536;
constexpr int kTuneReps = 3;
constexpr double kTuneTie = 1.03;
// Short-query classes: the ungapped pre-pass resolves most of their jobs, so their bins run few
// wavefronts and end as latency-bound tails of the batch; one common bin is worth more than a few
// percent per class, measured on all jobs of the class.
constexpr double kTuneTieShort = 1.10;

constexpr int kExtKs[] = {3, 5, 7, 9, 11, 13, 15, 17, 21, 25, 33};
constexpr int kNumExtKs = (int)(sizeof(kExtKs) / sizeof(kExtKs[0]));

int odd_k_index(int width) {
    for (int r = 0; r < kNumExtKs; ++r)
        if (16 * kExtKs[r] >= width) return r;
    return -1;
}

int odd_k_for(int width) {
    const int r = odd_k_index(width);
    return r < 0 ? -1 : kExtKs[r];
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

// The second half of binning: order the bins by stage then descending work, give them their
// offsets, and order the jobs by (bin, descending work, index) from each job's bin (its index in
// `bins` as created) and work key wk (65535 - work_key, so ascending wk is descending work).
// Small batches: one comparison sort; large ones: a stable LSD radix sort over the key
// (bin rank << 16 | wk), 12 bits a pass (two passes up to 256 bins), each pass split over `parts`
// threads (per-part digit histograms, offsets in (digit, part) order, a stable scatter), which
// gives the same order.
template <class BinOf>
HostVec<uint32_t> order_jobs(size_t n, std::vector<Item> &bins, BinOf bin_of, const uint16_t *wk, int parts) {
    std::vector<int> perm(bins.size());
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) {
        if (bins[a].stage != bins[b].stage) return bins[a].stage < bins[b].stage;
        return bins[a].work > bins[b].work;
    });
    std::vector<Item> sorted;
    std::vector<uint32_t> remap(bins.size());
    for (size_t r = 0; r < perm.size(); ++r) {
        remap[perm[r]] = (uint32_t)r;
        sorted.push_back(bins[perm[r]]);
    }
    uint32_t off = 0;
    for (auto &b : sorted) {
        b.offset = off;
        off += b.count;
    }
    bins.swap(sorted);
    HostVec<uint32_t> order(n);
    if (n < 16384) {
        std::iota(order.begin(), order.end(), 0u);
        std::sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
            const uint32_t bx = remap[bin_of(x)], by = remap[bin_of(y)];
            if (bx != by) return bx < by;
            if (wk[x] != wk[y]) return wk[x] < wk[y];
            return x < y;
        });
        return order;
    }
    // 12-bit digits: two passes for up to 256 bins, the first digit's histogram taken while the
    // keys are built and the last pass writing the order array itself
    constexpr int kDigit = 12, kRadix = 1 << kDigit;
    int key_bits = 16;
    while ((size_t)1 << (key_bits - 16) < bins.size()) ++key_bits;
    const int passes = (key_bits + kDigit - 1) / kDigit;
    HostVec<uint64_t> a(n), t(n);
    std::vector<uint32_t> hist((size_t)parts * kRadix);
    auto digit = [](uint64_t v, int pass) { return (uint32_t)(v >> (32 + kDigit * pass)) & (uint32_t)(kRadix - 1); };
    parallel_for(n, parts, [&](size_t lo, size_t hi, int q) {
        uint32_t *h = &hist[(size_t)q * kRadix];
        std::fill(h, h + kRadix, 0u);
        for (size_t k = lo; k < hi; ++k) {
            a[k] = ((uint64_t)((remap[bin_of(k)] << 16) | wk[k]) << 32) | (uint64_t)k;
            h[digit(a[k], 0)]++;
        }
    });
    for (int pass = 0; pass < passes; ++pass) {
        uint32_t run = 0;  // offsets in (digit, part) order: the scatter is stable
        for (int d = 0; d < kRadix; ++d)
            for (int q = 0; q < parts; ++q) {
                const uint32_t cnt = hist[(size_t)q * kRadix + d];
                hist[(size_t)q * kRadix + d] = run;
                run += cnt;
            }
        const bool last = pass + 1 == passes;
        parallel_for(n, parts, [&](size_t lo, size_t hi, int q) {
            uint32_t *h = &hist[(size_t)q * kRadix];
            if (last)
                for (size_t k = lo; k < hi; ++k) order[h[digit(a[k], pass)]++] = (uint32_t)a[k];
            else
                for (size_t k = lo; k < hi; ++k) t[h[digit(a[k], pass)]++] = a[k];
        });
        if (last) break;
        a.swap(t);
        parallel_for(n, parts, [&](size_t lo, size_t hi, int q) {
            uint32_t *h = &hist[(size_t)q * kRadix];
            std::fill(h, h + kRadix, 0u);
            for (size_t k = lo; k < hi; ++k) h[digit(a[k], pass + 1)]++;
        });
    }
    return order;
}

// Group jobs into bins by key (which may create bins), accumulate each bin's count and work, then
// order_jobs.
template <class KeyFn, class WorkFn>
HostVec<uint32_t> bin_jobs(size_t n, std::vector<Item> &bins, KeyFn key_of, WorkFn work_of) {
    HostVec<int> key(n);
    HostVec<uint16_t> wk(n);
    for (size_t k = 0; k < n; ++k) {
        key[k] = key_of(k);
        const uint64_t w = work_of(k);
        wk[k] = (uint16_t)(65535 - std::min<uint32_t>(work_key(w), 65535));
        bins[key[k]].count++;
        bins[key[k]].work += (double)w;
    }
    return order_jobs(n, bins, [&](size_t k) { return key[k]; }, wk.data(), 1);
}

// The bins of a plan, created on first use and found again through a table indexed by the kernel
// mapping (kind, K or column capacity, state width), not by a search per job.
struct BinSet {
    std::vector<Item> bins;
    int16_t lut[16][65][2];
    BinSet() { std::memset(lut, 0xff, sizeof lut); }
    int find(int kind, int param, bool narrow, int stage = 0) {
        int16_t &slot = lut[kind & 15][std::min(std::max(param, 0), 64)][narrow ? 1 : 0];
        if (slot >= 0) return slot;
        Item it;
        it.kind = kind;
        it.param = param;
        it.narrow = narrow;
        it.stage = stage;
        bins.push_back(it);
        slot = (int16_t)(bins.size() - 1);
        return slot;
    }
};

// Persistent grid of a bin: enough workgroups to fill every CU at the kernel's occupancy, but no
// more than the bin has work for.
int ext_grid(const bwg_ctx *c, const Tuning &tu, int kind, int param, bool narrow, int qcap, uint32_t count) {
    if (kind == kExtGroup) {
        const int per_cu = ext_group_blocks_per_cu(param, qcap, narrow);
        return (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)per_cu * c->num_cus * tu.ext_group_oversub,
                                                           ((int64_t)count + 15) / 16));
    }
    if (kind == kExtGroup8) {  // 32 groups a block
        const int per_cu = ext_group8_blocks_per_cu(param, qcap, narrow);
        return (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)per_cu * c->num_cus * tu.ext_group_oversub,
                                                           ((int64_t)count + 31) / 32));
    }
    const int per_cu = ext_thread_blocks_per_cu(param, narrow);
    return (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)per_cu * c->num_cus, ((int64_t)count + 63) / 64));
}

int ksw_grid(const bwg_ctx *c, const Tuning &tu, int K, bool rev, uint32_t count, bool packed = false) {
    const int per_cu = packed ? ksw_pk_blocks_per_cu(K, rev) : ksw_group_blocks_per_cu(K, rev);
    const int64_t groups = packed ? 256 / ksw_param_lanes(K) : 16;  // groups per block
    return (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)per_cu * c->num_cus * tu.ksw_group_oversub,
                                                       ((int64_t)count + groups - 1) / groups));
}

// Length class of an extension job (see bwg_tuner.h) and its minimal group-kernel K.
int ext_class_of(const ExtJob &j, const ExtScoring &sc, const Tuning &tu, int *kidx) {
    const int q = (int)j.qlen;
    const int64_t bound = (int64_t)j.h0 + (int64_t)q * std::max(sc.max_sc, 0);
    if (bound >= (1 << 18) || q >= 8192) return -1;
    // The group kernels' F scan adds up to 16 blocks of 33 columns' insertion extensions to exit
    // values below 2^18, in int32; a larger extension penalty goes to the plain kernel.
    if ((int64_t)sc.e_ins * 16 * 33 + (1 << 18) > INT32_MAX) return -1;
    const int w = clamp_band(sc.w, q, sc);
    *kidx = odd_k_index((int)std::min<int64_t>(q, 2 * (int64_t)w + 1));
    if (q > 32 && (*kidx < 0 || qcap_class(q) > tu.ext_group_max_qcap)) return -1;
    if (bound >= 65536 || tu.ext_force_wide) return -2;
    return q <= 32 ? (q - 1) / 4 : kExtShortClasses + *kidx;
}

// ext_class_of for a whole batch: everything but the h0 tests depends on the query length only
// (the band clamp takes two divisions), so it is computed once per length and looked up.
struct ExtClassifier {
    const ExtScoring &sc;
    const Tuning &tu;
    std::vector<int16_t> cls_of_q, kidx_of_q;  // class with a narrow-exact bound; -1 plain; -3 unset
    ExtClassifier(const ExtScoring &s, const Tuning &t) : sc(s), tu(t), cls_of_q(8192, -3), kidx_of_q(8192, -1) {}
    int classify(const ExtJob &j, int *kidx) {
        const int q = (int)j.qlen;
        if (q >= 8192) return ext_class_of(j, sc, tu, kidx);
        if (cls_of_q[q] == -3) {
            ExtJob probe = j;
            probe.h0 = 1;  // a bound below 65536 for every q < 8192 at scores below 8
            int ki = -1;
            int c = (int64_t)1 + (int64_t)q * std::max(sc.max_sc, 0) < 65536 ? ext_class_of(probe, sc, tu, &ki)
                                                                            : -4;
            if (c == -4) {  // large scores: no memo for this length
                return ext_class_of(j, sc, tu, kidx);
            }
            cls_of_q[q] = (int16_t)c;
            kidx_of_q[q] = (int16_t)ki;
        }
        *kidx = kidx_of_q[q];
        const int64_t bound = (int64_t)j.h0 + (int64_t)q * std::max(sc.max_sc, 0);
        const int c = cls_of_q[q];
        if (bound >= (1 << 18) || c == -1) return -1;
        if (bound >= 65536 || tu.ext_force_wide) return -2;
        return c;
    }
};

// Candidate mappings of a class, in the order the near-tie rule prefers: lanes per job (thread or
// group kernel), the group kernel's columns per lane (minimal K or the next), and the stored score
// width.
std::vector<TuneCandidate> ext_candidates(int cls) {
    std::vector<TuneCandidate> v;
    auto add = [&](int kind, int param, int width) {
        TuneCandidate t;
        t.kind = kind;
        t.param = param;
        t.width = width;
        v.push_back(t);
    };
    if (cls < kExtShortClasses) {
        // one mapping common to every short class first: the near-tie rule then gathers them into
        // one bin (after the pre-pass each holds few jobs, and every extra bin is a latency-bound
        // tail of the batch)
        const int qmax = 4 * cls + 4;
        add(kExtThread, 32, 16);
        if (qmax <= 16) add(kExtThread, 16, 16);
        add(kExtGroup, kExtKs[0], 32);
        add(kExtThread, qmax <= 16 ? 16 : 32, 32);
    } else {
        // 16-bit state before 32-bit state (twice the LDS a column): a near-tie between the two
        // goes to the 16-bit mapping, one K up where that is the one that ties
        const int r = cls - kExtShortClasses;
        add(kExtGroup, kExtKs[r], 16);
        if (r + 1 < kNumExtKs) add(kExtGroup, kExtKs[r + 1], 16);
        add(kExtGroup, kExtKs[r], 32);
        // eight jobs a wavefront, 2K + 1 columns per lane cover the same window
        if (kExtKs[r] <= 15) add(kExtGroup8, 2 * kExtKs[r] + 1, 16);  // 16-bit state: fits LDS at qcap 512
    }
    return v;
}

// Time each candidate mapping of one class on a sample of the batch's own jobs (already resident
// on the device) and record the fastest.  The sample kernels write the sampled jobs' true results,
// which the batch's own launches overwrite with the same values.
int tune_ext_class(bwg_ctx *c, bwg_batch *b, const ExtScoring &esc, uint64_t sig, int cls,
                   const std::vector<uint32_t> &members, hipStream_t st) {
    std::lock_guard<std::mutex> g(c->tuner.tune_mu);
    if (c->tuner.lookup(sig, cls, nullptr)) return BWG_OK;
    const uint32_t S = (uint32_t)std::min<size_t>(members.size(), kTuneSample);
    std::vector<uint32_t> idx(S);
    int qcap = 32;
    for (uint32_t r = 0; r < S; ++r) {
        idx[r] = members[(size_t)r * members.size() / S];
        qcap = std::max(qcap, qcap_class((int)b->hext[idx[r]].qlen));
    }
    TuneEntry ent;
    ent.cands = ext_candidates(cls);
    ent.sample = S;
    uint32_t *d = (uint32_t *)c->pool.alloc(((size_t)S + 1) * 4);
    if (!d) return BWG_ENOMEM;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMemcpyAsync(d + 1, idx.data(), (size_t)S * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    std::vector<double> best(ent.cands.size(), 1e30);
    for (int rep = 0; rep < kTuneReps && e == hipSuccess; ++rep) {
        for (size_t k = 0; k < ent.cands.size() && e == hipSuccess; ++k) {
            const TuneCandidate &cd = ent.cands[k];
            e = launch_zero_u32(st, d, 1);
            if (e == hipSuccess) e = hipEventRecord(e0, st);
            const bool narrow = cd.width == 16;
            const int grid = ext_grid(c, b->tu, cd.kind, cd.param, narrow, qcap, S);
            if (e == hipSuccess) {
                if (cd.kind == kExtGroup)
                    e = launch_ext_group(st, cd.param, narrow, b->d_arena, b->d_ref, (const ExtJob *)b->d_jobs, d + 1, S,
                                         nullptr, d, esc, qcap, (bwg_ext_result *)b->d_out, grid);
                else if (cd.kind == kExtGroup8)
                    e = launch_ext_group8(st, cd.param, narrow, b->d_arena, b->d_ref, (const ExtJob *)b->d_jobs, d + 1,
                                          S, nullptr, d, esc, qcap, (bwg_ext_result *)b->d_out, grid);
                else
                    e = launch_ext_thread(st, cd.param, narrow, b->d_arena, b->d_ref, (const ExtJob *)b->d_jobs, d + 1,
                                          S, nullptr, d, esc, (bwg_ext_result *)b->d_out, grid);
            }
            if (e == hipSuccess) e = hipEventRecord(e1, st);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float ms = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            if (e == hipSuccess) best[k] = std::min(best[k], (double)ms * 1e6 / S);
            c->launches += 2;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    c->pool.free(d);
    if (e != hipSuccess) return bwg_map_hip(e);
    // The first candidate (in the order of ext_candidates) within kTuneTie of the fastest:
    // near-ties are decided by timing noise, and the first candidates (the minimal K with 16-bit
    // state for window classes, one mapping common to every short class) need the fewest
    // registers and the least LDS and merge neighbouring classes into one bin, which the batch's
    // concurrent bins run better than extra bins.
    double fastest = 1e30;
    for (size_t k = 0; k < best.size(); ++k) {
        ent.cands[k].ns_per_job = best[k];
        fastest = std::min(fastest, best[k]);
    }
    const double tie = cls < kExtShortClasses ? kTuneTieShort : kTuneTie;
    for (size_t k = 0; k < best.size(); ++k)
        if (best[k] <= fastest * tie) {
            ent.choice = (int)k;
            break;
        }
    c->tuner.store(sig, cls, ent);
    return BWG_OK;
}

// Plan an extension batch at the scoring p->esc.  Round 0 may tune and runs the ungapped fast path
// first; a band-retry round (p->retry_round > 0) bins every job at its wider band, and its first
// item is the acceptance filter that fills the bins' compacted queues with the jobs sent back.
int plan_ext(bwg_ctx *c, bwg_batch *b, Plan *p, hipStream_t st) {
    const size_t n = b->n;
    const ExtScoring &sc = p->esc;
    const bool retry = p->retry_round > 0;
    const Tuning &tu = b->tu;
    BinSet bs;
    std::vector<Item> &bins = bs.bins;
    auto find = [&](int kind, int param, bool narrow) { return bs.find(kind, param, narrow); };
    // classify (over the host threads, each with its own memo), then tune the classes this batch
    // has enough jobs of and the table lacks
    const int parts = host_parts(tu, n);
    const int n_classes = kExtShortClasses + kNumExtKs;
    HostVec<int8_t> cls(n), kidx(n);
    std::vector<uint32_t> cls_count((size_t)parts * n_classes, 0u);
    parallel_for(n, parts, [&](size_t lo, size_t hi, int q) {
        ExtClassifier classifier(sc, tu);
        uint32_t *cc = &cls_count[(size_t)q * n_classes];
        for (size_t k = lo; k < hi; ++k) {
            int ki = -1;
            const int cl = classifier.classify(b->hext[k], &ki);
            cls[k] = (int8_t)cl;
            kidx[k] = (int8_t)ki;
            if (cl >= 0) cc[cl]++;
        }
    });
    const uint64_t sig = scoring_sig(sc);
    std::vector<int> choice_kind(n_classes, -1), choice_param(n_classes, 0), choice_width(n_classes, 16);
    std::vector<double> choice_ns(n_classes, 0.0);
    for (int cl = 0; cl < n_classes; ++cl) {
        if (!tu.autotune) break;
        size_t n_members = 0;
        for (int q = 0; q < parts; ++q) n_members += cls_count[(size_t)q * n_classes + cl];
        if (tu.tune_new && !retry && n_members >= kTuneMinJobs && !c->tuner.lookup(sig, cl, nullptr)) {
            std::vector<uint32_t> members;
            members.reserve(n_members);
            for (size_t k = 0; k < n; ++k)
                if (cls[k] == cl) members.push_back((uint32_t)k);
            int rc = wait_arena(b);  // the sample kernels read the arena
            if (!rc) rc = tune_ext_class(c, b, sc, sig, cl, members, st);
            if (rc) return rc;
        }
        TuneEntry ent;
        if (c->tuner.lookup(sig, cl, &ent)) {
            choice_kind[cl] = ent.cands[ent.choice].kind;
            choice_param[cl] = ent.cands[ent.choice].param;
            choice_width[cl] = ent.cands[ent.choice].width;
            choice_ns[cl] = ent.cands[ent.choice].ns_per_job;
        }
    }
    // Each job's kernel mapping as a code (kind << 8 | param << 1 | narrow), a pure function of its
    // class and length, so that the parts accumulate per-mapping counts, work and column capacity
    // on their own; the bins are then made in code order.
    auto code_of = [&](size_t k) -> uint32_t {
        const int q = (int)b->hext[k].qlen;
        const int cl = cls[k];
        int kind, param;
        bool narrow = cl >= 0;
        if (cl == -1) {
            kind = kExtPlain, param = 0, narrow = false;
        } else if (cl >= 0 && choice_kind[cl] >= 0) {
            kind = choice_kind[cl];
            param = choice_param[cl];
            narrow = choice_width[cl] == 16;
        } else if (q <= tu.ext_thread_max_q16) {
            kind = kExtThread, param = 16;
        } else if (q <= tu.ext_thread_max_q32) {
            kind = kExtThread, param = 32;
        } else {
            const int r = std::min(kidx[k] + (tu.ext_k_step ? 1 : 0), kNumExtKs - 1);
            kind = kExtGroup, param = kExtKs[r];
            if (tu.ext_group_lanes == 8 && param <= 15 &&
                ext_group8_lds_bytes(2 * param + 1, qcap_class(q), narrow) <= kMaxBlockLds)
                kind = kExtGroup8, param = 2 * param + 1;
        }
        return ((uint32_t)kind << 8) | ((uint32_t)param << 1) | (narrow ? 1u : 0u);
    };
    constexpr int kCodes = 16 << 8;
    struct CodeAcc {
        uint32_t count = 0;
        int32_t qcap = 0;
        double work = 0.0;
        double est_ns = 0.0;
        uint32_t untimed = 0;  // jobs without a tuned class
    };
    std::vector<CodeAcc> acc((size_t)parts * kCodes);
    HostVec<uint16_t> code(n), wk(n);
    parallel_for(n, parts, [&](size_t lo, size_t hi, int q) {
        CodeAcc *a = &acc[(size_t)q * kCodes];
        for (size_t k = lo; k < hi; ++k) {
            const uint32_t cd = code_of(k);
            const ExtJob &j = b->hext[k];
            const uint64_t w = (uint64_t)j.qlen * (j.tlen + 1);
            code[k] = (uint16_t)cd;
            wk[k] = (uint16_t)(65535 - std::min<uint32_t>(work_key(w), 65535));
            a[cd].count++;
            a[cd].work += (double)w;
            const int cl = cls[k];
            if (cl >= 0 && choice_ns[cl] > 0.0) a[cd].est_ns += choice_ns[cl];
            else a[cd].untimed++;
            if (is_ext_group((int)(cd >> 8))) a[cd].qcap = std::max(a[cd].qcap, qcap_class((int)j.qlen));
        }
    });
    std::vector<int16_t> bin_of_code(kCodes, -1);
    for (int cd = 0; cd < kCodes; ++cd) {
        CodeAcc t;
        for (int q = 0; q < parts; ++q) {
            const CodeAcc &a = acc[(size_t)q * kCodes + cd];
            t.count += a.count;
            t.work += a.work;
            t.est_ns += a.est_ns;
            t.untimed += a.untimed;
            t.qcap = std::max(t.qcap, a.qcap);
        }
        if (!t.count) continue;
        const int kind = cd >> 8;
        const int bi = find(kind, (cd >> 1) & 127, (cd & 1) != 0);
        bins[bi].count += t.count;
        bins[bi].work += t.work;
        bins[bi].est_ns += t.est_ns;
        bins[bi].est_all = (bins[bi].count == t.count) ? t.untimed == 0 : bins[bi].est_all && t.untimed == 0;
        if (is_ext_group(kind)) bins[bi].qcap = std::max(bins[bi].qcap, t.qcap);
        bin_of_code[cd] = (int16_t)bi;
    }
    p->order = order_jobs(n, bins, [&](size_t k) { return (uint32_t)bin_of_code[code[k]]; }, wk.data(), parts);
    for (size_t r = 0; r < bins.size(); ++r) {
        Item &it = bins[r];
        it.head = (uint32_t)r;
        if (is_ext_dp(it.kind)) {
            it.grid = ext_grid(c, tu, it.kind, it.param, it.narrow, it.qcap, it.count);
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
            p->d_plain_jobs = batch_alloc(c, b, std::max<size_t>(1, it.count) * sizeof(ExtJob));
            p->d_plain_scratch = batch_alloc(c, b, std::max<size_t>(1, total) * sizeof(int2));
            if (!p->d_plain_jobs || !p->d_plain_scratch) return BWG_ENOMEM;
        }
    }
    p->n_heads = std::max<int>(1, (int)bins.size());
    // The reference's ungapped-diagonal fast path (crates/bwa-neon/src/batched.rs:716-798) applies
    // to uniform DNA matrices: run it first over the group and thread bins, which then drain the
    // compacted queues of the jobs it declines.  A band-retry round drains the queues its
    // acceptance filter fills instead.
    const int32_t a = sc.mat[0], bmis = -(int32_t)sc.mat[1];
    const int32_t xmax = ungapped_xmax(sc);
    p->ungapped = !retry && tu.ungapped && is_uniform_dna(sc) && a > 0 && bmis > 0 && xmax >= 0 && !bins.empty();
    p->compact = (p->ungapped || retry) && !bins.empty();
    if (p->compact) {
        p->n_bins = (int)bins.size();
        // [heads nb][counters, kCountStride apart][slot offsets nb][skip flags nb]
        p->meta.assign((size_t)p->n_bins * (kCountStride + 3), 0u);
        bool any = false;
        for (int r = 0; r < p->n_bins; ++r) {
            p->meta[seg_word(p) + r] = bins[r].offset;
            const bool dp = is_ext_dp(bins[r].kind);
            p->meta[seg_word(p) + p->n_bins + r] = dp ? 0u : 1u;
            any |= dp;
        }
        if (p->ungapped) p->ungapped = any && p->n_bins <= 255;
        p->compact = p->ungapped || retry;
    }
    if (p->compact) {
        for (auto &it : bins) it.stage = 1;
        Item u;
        u.stage = 0;
        u.count = (uint32_t)n;
        if (p->ungapped) {
            p->ungapped_xmax = xmax;
            u.kind = kExtUngapped;
            u.grid = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)c->num_cus * 8, ((int64_t)n + 15) / 16));
            p->d_declined = (uint8_t *)batch_alloc(c, b, std::max<size_t>(1, n));
            if (!p->d_declined) return BWG_ENOMEM;
        } else {
            u.kind = kExtRetryFilter;
        }
        bins.insert(bins.begin(), u);
        p->d_corder = (uint32_t *)batch_alloc(c, b, std::max<size_t>(1, n) * 4);
        if (!p->d_corder) return BWG_ENOMEM;
    }
    p->items = bins;
    p->d_order = (uint32_t *)batch_alloc(c, b, std::max<size_t>(1, n) * 4);
    p->d_heads = (uint32_t *)batch_alloc(c, b, std::max<size_t>((size_t)p->n_heads, p->meta.size()) * 4);
    if (!p->d_order || !p->d_heads) return BWG_ENOMEM;
    hipError_t e = hipMemcpyAsync(p->d_order, p->order.data(), n * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && p->compact)
        e = hipMemcpyAsync(p->d_heads, p->meta.data(), p->meta.size() * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && !p->ext_plain_jobs.empty())
        e = hipMemcpyAsync(p->d_plain_jobs, p->ext_plain_jobs.data(), p->ext_plain_jobs.size() * sizeof(ExtJob),
                           hipMemcpyHostToDevice, st);
    return bwg_map_hip(e);
}

// Time the rescue forward kernel at K and at the next K up on a sample of one class's jobs.  The
// sample runs on its own descriptor copies (row-maximum slots tiled as in the batch plan) and writes
// partial results to the first slots of the batch's output, which the batch's own launches
// overwrite.
int tune_ksw_class(bwg_ctx *c, bwg_batch *b, uint64_t sig, int cls, const std::vector<uint32_t> &members,
                   hipStream_t st) {
    std::lock_guard<std::mutex> g(c->tuner.tune_mu);
    if (c->tuner.lookup(sig, cls, nullptr)) return BWG_OK;
    const uint32_t S = (uint32_t)std::min<size_t>(members.size(), kTuneSample / 2);
    std::vector<KswJob> jobs(S);
    std::vector<uint32_t> ident(S);
    uint64_t rows = 0;
    for (uint32_t t0 = 0; t0 < S; t0 += 64) {
        const uint32_t t1 = std::min<uint32_t>(S, t0 + 64);
        uint32_t tmax = 0;
        for (uint32_t r = t0; r < t1; ++r) {
            jobs[r] = b->hksw[members[(size_t)r * members.size() / S]];
            jobs[r].aux0 = (uint32_t)(rows >> 6);
            jobs[r].aux1 = r - t0;
            ident[r] = r;
            tmax = std::max(tmax, jobs[r].tlen);
        }
        rows += 64ull * tmax;
    }
    TuneEntry ent;
    ent.sample = S;
    const int K = 2 * (cls % 8 + 1);
    for (int k2 = K; k2 <= std::min(16, K + 2); k2 += 2) {
        TuneCandidate t;
        t.kind = kKswFwd;
        t.param = k2;
        ent.cands.push_back(t);
    }
    if (cls < 8 && !b->tu.ksw_force_wide) {  // 16-bit-exact jobs: the packed kernel too
        TuneCandidate t;
        t.kind = kKswFwd;
        t.param = K;
        t.width = 16;
        ent.cands.push_back(t);
        if (ksw_pk8_scoring_ok(b->ksc)) {  // and with eight jobs a wavefront
            t.param = ksw_param(K, 8);
            ent.cands.push_back(t);
        }
    }
    uint32_t *d = (uint32_t *)c->pool.alloc(((size_t)S + 1) * 4);
    KswJob *dj = (KswJob *)c->pool.alloc((size_t)S * sizeof(KswJob));
    int32_t *drm = (int32_t *)c->pool.alloc(std::max<uint64_t>(1, rows) * 4);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = (d && dj && drm) ? hipSuccess : hipErrorOutOfMemory;
    if (e == hipSuccess) e = hipMemcpyAsync(d + 1, ident.data(), (size_t)S * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dj, jobs.data(), (size_t)S * sizeof(KswJob), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    std::vector<double> best(ent.cands.size(), 1e30);
    for (int rep = 0; rep < kTuneReps && e == hipSuccess; ++rep) {
        for (size_t k = 0; k < ent.cands.size() && e == hipSuccess; ++k) {
            const int kk = ent.cands[k].param;
            const bool packed = ent.cands[k].width == 16;
            e = launch_zero_u32(st, d, 1);
            if (e == hipSuccess) e = hipEventRecord(e0, st);
            if (e == hipSuccess) {
                if (packed)
                    e = launch_ksw_pk(st, kk, false, b->d_arena, b->d_ref, dj, d + 1, S, d, b->ksc,
                                      (bwg_ksw_result *)b->d_out, drm, ksw_grid(c, b->tu, kk, false, S, true));
                else
                    e = launch_ksw_group(st, kk, false, b->d_arena, b->d_ref, dj, d + 1, S, d, b->ksc,
                                         (bwg_ksw_result *)b->d_out, drm, ksw_grid(c, b->tu, kk, false, S));
            }
            if (e == hipSuccess) e = hipEventRecord(e1, st);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float ms = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            if (e == hipSuccess) best[k] = std::min(best[k], (double)ms * 1e6 / S);
            c->launches += 2;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    c->pool.free(d);
    c->pool.free(dj);
    c->pool.free(drm);
    if (e != hipSuccess) return bwg_map_hip(e);
    for (size_t k = 0; k < best.size(); ++k) {
        ent.cands[k].ns_per_job = best[k];
        if (best[k] < best[ent.choice]) ent.choice = (int)k;
    }
    c->tuner.store(sig, cls, ent);
    return BWG_OK;
}

int plan_ksw(bwg_ctx *c, bwg_batch *b, Plan *p, hipStream_t st) {
    const size_t n = b->n;
    const KswScoring &sc = b->ksc;
    const Tuning &tu = b->tu;
    BinSet bs;
    std::vector<Item> &bins = bs.bins;
    // fast bins first in the order array
    auto find = [&](int kind, int param, bool packed) { return bs.find(kind, param, packed, kind == kKswFwd ? 0 : 1); };
    // class = the minimal K (index K/2-1), +8 for jobs whose scores may not fit 16 bits; tune the
    // classes with enough jobs
    const bool pk_scoring = ksw_pk_scoring_ok(sc);
    const int32_t mm = std::max<int32_t>(1, mat_max(sc.mat));
    std::vector<int8_t> cls_of(n);
    std::vector<std::vector<uint32_t>> members(16);
    for (size_t k = 0; k < n; ++k) {
        const int qlen = (int)b->hksw[k].qlen;
        const int K = sc.o_ins >= 0 && b->hksw[k].tlen <= kMaxFastTlen ? ksw_k_for(widest_pad(qlen, sc.max_sc)) : -1;
        const bool pk = pk_scoring && (int64_t)qlen * mm < kKswPkMaxScore;
        cls_of[k] = (int8_t)(K < 0 ? -1 : K / 2 - 1 + (pk ? 0 : 8));
        if (K > 0) members[cls_of[k]].push_back((uint32_t)k);
    }
    const uint64_t sig = scoring_sig(sc);
    std::vector<int> chosen(16, 0);
    std::vector<char> chosen_pk(16, 0);
    for (int cl = 0; cl < 16; ++cl) {
        const int K = 2 * (cl % 8 + 1);
        chosen[cl] = tu.ksw_k_step ? std::min(16, K + 2) : K;
        chosen_pk[cl] = cl < 8 && !tu.ksw_force_wide;
        if (chosen_pk[cl] && tu.ksw_group_lanes == 8 && ksw_pk8_scoring_ok(sc)) chosen[cl] = ksw_param(chosen[cl], 8);
        if (!tu.autotune) continue;
        // a rescue-window batch's descriptors are written on the device, so it only reads the table
        if (tu.tune_new && !b->matesw && members[cl].size() >= kTuneMinJobs && !c->tuner.lookup(sig, cl, nullptr)) {
            int rc = wait_arena(b);  // the sample kernels read the arena
            if (!rc) rc = tune_ksw_class(c, b, sig, cl, members[cl], st);
            if (rc) return rc;
        }
        TuneEntry ent;
        if (c->tuner.lookup(sig, cl, &ent)) {
            chosen[cl] = ent.cands[ent.choice].param;
            chosen_pk[cl] = ent.cands[ent.choice].width == 16;
        }
    }
    auto key_of = [&](size_t k) {
        const int cl = cls_of[k];
        return cl < 0 ? find(kKswPlain, 0, false) : find(kKswFwd, chosen[cl], chosen_pk[cl] != 0);
    };
    auto work_of = [&](size_t k) { return (uint64_t)b->hksw[k].tlen; };
    p->order = bin_jobs(n, bins, key_of, work_of);
    // Row-maximum slots for the fast-path jobs: tiles of 64 consecutive jobs of the order array,
    // row-major within a tile ([row][64]), so the tracker kernel's 64 lanes read consecutive words.
    // A job's slot for row i is row_slot(job) + 64*i.
    p->ksw_jobs.assign(b->hksw.begin(), b->hksw.end());
    for (auto &it : bins)
        if (it.kind == kKswFwd) p->n_fast += it.count;
    uint64_t rows = 0;
    for (uint32_t t0 = 0; t0 < p->n_fast; t0 += 64) {
        uint32_t tmax = 0;
        const uint32_t t1 = std::min<uint32_t>(p->n_fast, t0 + 64);
        for (uint32_t r = t0; r < t1; ++r) tmax = std::max(tmax, p->ksw_jobs[p->order[r]].tlen);
        for (uint32_t r = t0; r < t1; ++r) {
            p->ksw_jobs[p->order[r]].aux0 = (uint32_t)(rows >> 6);
            p->ksw_jobs[p->order[r]].aux1 = r - t0;
        }
        rows += 64ull * tmax;
    }
    std::vector<Item> items;
    int heads = 0;
    for (auto &it : bins) {
        if (it.kind == kKswFwd) {
            Item f = it;
            f.stage = 0;
            f.head = (uint32_t)heads++;
            f.grid = ksw_grid(c, tu, it.param, false, it.count, it.narrow);
            items.push_back(f);
            Item r = it;
            r.kind = kKswRev;
            r.stage = 2;
            r.head = (uint32_t)heads++;
            r.grid = ksw_grid(c, tu, it.param, true, it.count, it.narrow);
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
    std::vector<int32_t> plain_slot;
    if (b->matesw) {
        // the windows (and the mates' reverse complements) are computed on the device first
        for (auto &it : items) it.stage += 1;
        Item rc;
        rc.kind = kKswMateRc;
        rc.count = b->n_anchors;
        items.push_back(rc);
        Item wi;
        wi.kind = kKswWindows;
        wi.count = b->n_anchors;
        items.push_back(wi);
        if (!p->ksw_plain_jobs.empty()) {
            plain_slot.assign(n, -1);
            for (const auto &it : bins)
                if (it.kind == kKswPlain)
                    for (uint32_t r2 = 0; r2 < it.count; ++r2) plain_slot[p->order[it.offset + r2]] = (int32_t)r2;
            p->d_plain_slot = (int32_t *)c->pool.alloc(n * 4);
            if (!p->d_plain_slot) return BWG_ENOMEM;
        }
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
    if (e == hipSuccess && !plain_slot.empty())
        e = hipMemcpyAsync(p->d_plain_slot, plain_slot.data(), n * 4, hipMemcpyHostToDevice, st);
    return bwg_map_hip(e);
}

// Global alignment with traceback: bins by the group kernel's K for the band's widest row
// (min(qlen, 2w+1) columns), each job's traceback codes and CIGAR scratch laid out by the plan.
int plan_glb(bwg_ctx *c, bwg_batch *b, Plan *p, hipStream_t st) {
    const size_t n = b->n;
    std::vector<Item> bins;
    std::vector<GlobalJob> jobs = b->hglb;
    // queries beyond the group kernel's LDS columns, or bands wider than its widest K, run one job
    // per thread (k = 0 marks them for the traceback kernel, which skips them)
    auto key_of = [&](size_t k) {
        GlobalJob &j = jobs[k];
        const int64_t ncol = std::min<int64_t>(j.qlen, 2 * (int64_t)std::max(j.w, 0) + 1);
        const int K = j.qlen > kGlbGroupMaxQ ? -1 : odd_k_for((int)std::max<int64_t>(ncol, 1));
        j.k = K < 0 ? 0u : (uint32_t)K;
        const int kind = K < 0 ? kGlbPlain : kGlbGroup;
        const int param = K < 0 ? 0 : K;
        for (size_t r = 0; r < bins.size(); ++r)
            if (bins[r].kind == kind && bins[r].param == param) {
                bins[r].qcap = std::max(bins[r].qcap, qcap_class((int)j.qlen));
                return (int)r;
            }
        Item it;
        it.kind = kind;
        it.param = param;
        it.qcap = qcap_class((int)j.qlen);
        bins.push_back(it);
        return (int)bins.size() - 1;
    };
    auto work_of = [&](size_t k) { return (uint64_t)jobs[k].tlen * (std::min<uint64_t>(jobs[k].qlen, 64) + 1); };
    p->order = bin_jobs(n, bins, key_of, work_of);
    uint64_t zw = 0, ow = 0;
    for (size_t k = 0; k < n; ++k) {
        GlobalJob &j = jobs[k];
        if (j.k == 0) zw = (zw + 1) & ~1ull;  // its {H, E} pairs 8-byte aligned
        j.z_off = zw;
        j.ops_off = ow;
        if (j.k == 0) {  // one job per thread: {H, E} columns, then a code byte per band cell
            const uint64_t ncol = std::min<uint64_t>(j.qlen, 2 * (uint64_t)std::max(j.w, 0) + 1);
            zw += 2 * ((uint64_t)j.qlen + 2) + ((uint64_t)j.tlen * ncol + 3) / 4;
        } else {
            zw += (uint64_t)((j.tlen + 7) & ~7u) * 16 * ((j.k + 7) / 8);  // rows in tiles of 8
        }
        ow += (uint64_t)j.qlen + j.tlen + 2;
    }
    std::vector<Item> items;
    for (size_t r = 0; r < bins.size(); ++r) {
        Item it = bins[r];
        it.head = (uint32_t)r;
        it.stage = 0;
        if (it.kind == kGlbPlain) {
            it.grid = (int)std::max<int64_t>(1, ((int64_t)it.count + 63) / 64);
        } else {
            it.narrow = glb_scores_packable(b->gsc);  // the packed-score kernel form
            const int per_cu = glb_group_blocks_per_cu(it.param, it.qcap, it.narrow);
            it.grid = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)per_cu * c->num_cus,
                                                               ((int64_t)it.count + 15) / 16));
        }
        items.push_back(it);
    }
    Item tr;
    tr.kind = kGlbTrace;
    tr.stage = 1;
    tr.count = (uint32_t)n;
    items.push_back(tr);
    p->items = items;
    p->n_heads = std::max<int>(1, (int)bins.size());
    p->d_order = (uint32_t *)c->pool.alloc(std::max<size_t>(1, n) * 4);
    p->d_heads = (uint32_t *)c->pool.alloc((size_t)p->n_heads * 4);
    p->d_jobs = c->pool.alloc(std::max<size_t>(1, n) * sizeof(GlobalJob));
    b->d_z = (uint32_t *)c->pool.alloc(std::max<uint64_t>(1, zw) * 4);
    b->d_ops = (uint32_t *)c->pool.alloc(std::max<uint64_t>(1, ow) * 4);
    b->d_nops = (uint32_t *)c->pool.alloc(std::max<size_t>(1, n) * 4);
    if (!p->d_order || !p->d_heads || !p->d_jobs || !b->d_z || !b->d_ops || !b->d_nops) return BWG_ENOMEM;
    hipError_t e = hipMemcpyAsync(p->d_order, p->order.data(), n * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(p->d_jobs, jobs.data(), n * sizeof(GlobalJob), hipMemcpyHostToDevice, st);
    b->hglb = jobs;  // with the bins' K and the layout offsets, for the CIGAR gather
    return bwg_map_hip(e);
}

hipError_t launch_item(bwg_batch *b, Plan *p, const Item &it, hipStream_t st) {
    switch (it.kind) {
    case kExtGroup:
    case kExtGroup8:
    case kExtThread: {
        // after the ungapped pre-pass or a retry filter the bin drains its compacted queue
        const uint32_t *order = p->compact ? p->d_corder + it.offset : p->d_order + it.offset;
        const uint32_t *dcount = p->compact ? p->d_heads + p->n_bins + kCountStride * it.head : nullptr;
        if (it.kind == kExtGroup)
            return launch_ext_group(st, it.param, it.narrow, b->d_arena, b->d_ref, (const ExtJob *)b->d_jobs, order, it.count,
                                    dcount, p->d_heads + it.head, p->esc, it.qcap, (bwg_ext_result *)b->d_out,
                                    it.grid);
        if (it.kind == kExtGroup8)
            return launch_ext_group8(st, it.param, it.narrow, b->d_arena, b->d_ref, (const ExtJob *)b->d_jobs, order,
                                     it.count, dcount, p->d_heads + it.head, p->esc, it.qcap,
                                     (bwg_ext_result *)b->d_out, it.grid);
        return launch_ext_thread(st, it.param, it.narrow, b->d_arena, b->d_ref, (const ExtJob *)b->d_jobs, order,
                                 it.count, dcount, p->d_heads + it.head, p->esc, (bwg_ext_result *)b->d_out, it.grid);
    }
    case kExtRetryFilter: {
        const int r = p->retry_round;
        const int32_t w_prev = b->esc.w << (r - 1), w_next = b->esc.w << r;
        return launch_ext_retry_filter(st, p->d_order, it.count, p->d_heads + seg_word(p),
                                       p->d_heads + seg_word(p) + p->n_bins, p->n_bins,
                                       (const bwg_ext_result *)b->d_out, b->d_prev0, b->d_prev, b->d_active, r == 1,
                                       w_prev, w_next, b->d_wout, p->d_corder, p->d_heads + p->n_bins);
    }
    case kExtUngapped:
        return launch_ext_ungapped(st, b->d_arena, b->d_ref, (const ExtJob *)b->d_jobs, p->d_order, it.count,
                                   p->d_heads + seg_word(p), p->d_heads + seg_word(p) + p->n_bins, p->n_bins,
                                   p->d_declined, p->d_corder, p->d_heads + p->n_bins, p->esc, p->ungapped_xmax,
                                   (bwg_ext_result *)b->d_out, it.grid, p->d_heads, (uint32_t)seg_word(p));
    case kExtPlain:
        return launch_ext_plain_scatter(st, b->d_arena, b->d_ref, (const ExtJob *)p->d_plain_jobs,
                                        p->d_order + it.offset, it.count, p->esc, (int2 *)p->d_plain_scratch,
                                        (bwg_ext_result *)b->d_out, p->retry_round > 0 ? b->d_active : nullptr);
    case kKswFwd:
    case kKswRev:
        if (it.narrow)
            return launch_ksw_pk(st, it.param, it.kind == kKswRev, b->d_arena, b->d_ref, (const KswJob *)p->d_jobs,
                                 p->d_order + it.offset, it.count, p->d_heads + it.head, b->ksc,
                                 (bwg_ksw_result *)b->d_out, p->d_rowmax, it.grid);
        return launch_ksw_group(st, it.param, it.kind == kKswRev, b->d_arena, b->d_ref, (const KswJob *)p->d_jobs,
                                p->d_order + it.offset, it.count, p->d_heads + it.head, b->ksc,
                                (bwg_ksw_result *)b->d_out, p->d_rowmax, it.grid);
    case kGlbGroup:
        return launch_glb_group(st, it.param, it.narrow, b->d_arena, b->d_ref, (const GlobalJob *)p->d_jobs,
                                p->d_order + it.offset, it.count, p->d_heads + it.head, b->gsc, it.qcap, b->d_z,
                                (int32_t *)b->d_out, it.grid);
    case kGlbPlain:
        return launch_glb_plain(st, b->d_arena, b->d_ref, (const GlobalJob *)p->d_jobs, p->d_order + it.offset,
                                it.count, b->gsc, b->d_z, b->d_ops, b->d_nops, (int32_t *)b->d_out);
    case kGlbTrace:
        return launch_glb_traceback(st, (const GlobalJob *)p->d_jobs, it.count, b->d_z, b->d_ops, b->d_nops);
    case kKswMateRc:
        return launch_mate_revcomp(st, b->d_arena, b->d_anchors, b->n_anchors);
    case kKswWindows:
        return launch_matesw_windows(st, b->d_anchors, b->n_anchors, b->d_ma_rb, b->mp, b->d_contig_off,
                                     b->d_contig_len, (KswJob *)p->d_jobs, p->d_plain_slot,
                                     (KswJob *)p->d_plain_jobs, b->d_win);
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
    // queue heads (and, with the ungapped pre-pass, the compacted queue lengths) start at zero:
    // cleared by the pre-pass's first block where the sequence starts with it, else by a fill kernel
    const bool prepass_first = !p->items.empty() && p->items[0].kind == kExtUngapped;
    hipError_t e = prepass_first
                       ? hipSuccess
                       : launch_zero_u32(ss.main, p->d_heads, (uint32_t)(p->compact ? seg_word(p) : p->n_heads));
    int64_t nl = prepass_first ? 0 : 1;
    size_t i = 0;
    while (i < p->items.size() && e == hipSuccess) {
        size_t j = i;
        while (j < p->items.size() && p->items[j].stage == p->items[i].stage) ++j;
        const int nitems = (int)(j - i);
        const bool par = b->tu.concurrent_bins && nitems > 1;
        if (par) {
            e = hipEventRecord(ss.fork, ss.main);
            for (int k = 0; k < kAuxStreams && e == hipSuccess; ++k) e = hipStreamWaitEvent(ss.aux[k], ss.fork, 0);
        }
        // largest expected work first, each onto the stream with the least work queued
        // (longest-processing-time-first over the stage's items).  A bin that drains a compacted
        // queue runs only the jobs the pre-pass (or the retry filter) passes on: its nominal work is
        // weighted by the share of its jobs measured on this batch's previous run, else on earlier
        // batches of the context.  Short-query bins lose most of their jobs to the pre-pass; counted
        // at their nominal work, they were dealt streams ahead of long bins.
        // When every DP bin of the stage has a time per job from the autotuner for each of its jobs,
        // the bins are dealt by that time (the same share applied) rather than by cells: a bin of
        // short, narrow rows costs more per cell than one of long, wide rows.
        // Other items of such a stage (the plain bin) are converted at the timed bins' mean time
        // per nominal cell.
        bool timed_stage = true;
        double t_ns = 0.0, t_work = 0.0;
        for (int k = 0; k < nitems; ++k) {
            const Item &it = p->items[i + k];
            if (!is_ext_dp(it.kind)) continue;
            if (!(it.est_all && it.est_ns > 0)) timed_stage = false;
            t_ns += it.est_ns;
            t_work += it.work;
        }
        const double ns_per_cell = t_work > 0 ? t_ns / t_work : 0.0;
        std::vector<double> wexp(nitems), shares(nitems, -1.0);
        for (int k = 0; k < nitems; ++k) {
            const Item &it = p->items[i + k];
            double share = 1.0;
            if (p->compact && is_ext_dp(it.kind)) {
                if (it.dp_share >= 0) {
                    share = it.dp_share;
                } else {
                    std::lock_guard<std::mutex> g(c->dp_mu);
                    const float v = c->dp_share[item_code(it)];
                    if (v >= 0) share = v;
                }
                if (it.dp_share >= 0 || share < 1.0) shares[k] = share;
            }
            wexp[k] = (!timed_stage ? it.work : is_ext_dp(it.kind) ? it.est_ns : it.work * ns_per_cell) * share;
        }
        std::vector<size_t> ord(nitems);
        std::iota(ord.begin(), ord.end(), i);
        std::stable_sort(ord.begin(), ord.end(), [&](size_t x, size_t y) { return wexp[x - i] > wexp[y - i]; });
        double load[kAuxStreams + 1] = {};
        const bool timed = b->tu.item_timing;
        if (timed && p->ev0.size() != p->items.size()) {
            p->ev0.assign(p->items.size(), nullptr);
            p->ev1.assign(p->items.size(), nullptr);
            for (size_t k = 0; k < p->items.size() && e == hipSuccess; ++k) {
                e = hipEventCreate(&p->ev0[k]);
                if (e == hipSuccess) e = hipEventCreate(&p->ev1[k]);
            }
        }
        // A one-job-per-lane bin draining a compacted queue holds few of its nominal jobs: its grid
        // is sized for the declined share measured on an earlier run (the persistent loop takes any
        // surplus from the queue counter), so that its one-wave blocks do not hold CU slots the
        // long bins' grids use.  Such a bin lasts as long as its longest lanes' jobs whatever its
        // share of the GPU, so it is not dealt by its (small) work: it goes after the stage's
        // largest bin, on that bin's stream.  That bin's blocks fill the CUs while it runs and it
        // ends well before the bins that share the GPU with it; dealt by work, the short bin went
        // last onto some other stream and ran alone at the batch's end.
        auto latency_bin = [&](size_t r) { return par && p->compact && p->items[r].kind == kExtThread; };
        int first_slot = -1;
        std::vector<size_t> deferred;
        auto issue = [&](size_t r, int slot) {
            hipStream_t st = (!par || slot == 0) ? ss.main : ss.aux[slot - 1];
            Item it = p->items[r];
            if (p->compact && it.kind == kExtThread && shares[r - i] >= 0) {
                const double waves = std::ceil((double)it.count * shares[r - i] * 1.5 / 64.0) + 16.0;
                it.grid = (int)std::max(1.0, std::min((double)it.grid, waves));
            }
            p->items[r].launch_grid = it.grid;
            if (timed) e = hipEventRecord(p->ev0[r], st);
            if (e == hipSuccess) e = launch_item(b, p, it, st);
            if (timed && e == hipSuccess) e = hipEventRecord(p->ev1[r], st);
            ++nl;
        };
        // a stage of such bins alone is dealt by work like any other
        bool others = false;
        for (size_t r : ord) others |= !latency_bin(r);
        for (size_t r : ord) {
            if (e != hipSuccess) break;
            if (others && latency_bin(r)) {
                deferred.push_back(r);
                continue;
            }
            int slot = 0;
            for (int k = 1; k < std::min(b->tu.bin_streams, kAuxStreams + 1); ++k)
                if (load[k] < load[slot]) slot = k;
            load[slot] += std::max(wexp[r - i], 1.0);
            if (first_slot < 0) first_slot = slot;
            issue(r, slot);
        }
        for (size_t r : deferred)
            if (e == hipSuccess) issue(r, first_slot);
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

// Plans of a batch: one, or for a band-retry batch one per round, the round-r plan binning every job
// at band w << r.
int plan_batch(bwg_ctx *c, bwg_batch *b, hipStream_t st) {
    if (b->plan) return BWG_OK;
    Plan *p = new Plan;
    b->plan = p;
    p->esc = b->esc;
    int rc = b->kind == 0 ? plan_ext(c, b, p, st) : b->kind == 1 ? plan_ksw(c, b, p, st) : plan_glb(c, b, p, st);
    for (int r = 1; r < b->side_tries && !rc; ++r) {
        Plan *q = new Plan;
        b->retry_plans.push_back(q);
        q->esc = b->esc;
        q->esc.w = b->esc.w << r;
        q->retry_round = r;
        rc = plan_ext(c, b, q, st);
    }
    return rc;
}

// The whole launch sequence of a batch: its plan, then each retry round's (filter, then bins).  With
// a single band try every job keeps the first band.
hipError_t emit_all(bwg_ctx *c, bwg_batch *b, StreamSet &ss, int64_t *launches) {
    int64_t nl = 0;
    hipError_t e = emit(c, b, (Plan *)b->plan, ss, &nl);
    for (void *q : b->retry_plans) {
        int64_t k = 0;
        if (e == hipSuccess) e = emit(c, b, (Plan *)q, ss, &k);
        nl += k;
    }
    if (e == hipSuccess && b->side_tries == 1) {
        e = launch_fill_i32(ss.main, b->d_wout, (uint32_t)b->n, b->esc.w);
        ++nl;
    }
    *launches = nl;
    return e;
}

// The compacted-queue lengths the last run left on the device (each bin's jobs that the ungapped
// pre-pass or the retry filter passed to the DP), as each item's share of its jobs, recorded in the
// plan and in the context's table for later batches.  Reads after the last run in stream order.
hipError_t learn_plan_shares(bwg_ctx *c, Plan *p, hipStream_t st) {
    if (!p->compact || p->n_bins == 0) return hipSuccess;
    std::vector<uint32_t> cnt((size_t)kCountStride * p->n_bins);
    hipError_t e = hipMemcpyAsync(cnt.data(), p->d_heads + p->n_bins, cnt.size() * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> g(c->dp_mu);
    for (Item &it : p->items)
        if (is_ext_dp(it.kind) && it.count > 0) {
            it.dp_share = std::min(1.0, (double)cnt[(size_t)kCountStride * it.head] / (double)it.count);
            if (p->retry_round == 0) c->dp_share[item_code(it)] = (float)it.dp_share;
        }
    return hipSuccess;
}

hipError_t learn_dp_shares(bwg_ctx *c, bwg_batch *b, hipStream_t st) {
    hipError_t e = b->plan ? learn_plan_shares(c, (Plan *)b->plan, st) : hipSuccess;
    for (void *q : b->retry_plans)
        if (e == hipSuccess) e = learn_plan_shares(c, (Plan *)q, st);
    return e;
}

void free_plan(bwg_ctx *c, Plan *p) {
    for (size_t k = 0; k < p->ev0.size(); ++k) {
        if (p->ev0[k]) (void)hipEventDestroy(p->ev0[k]);
        if (p->ev1[k]) (void)hipEventDestroy(p->ev1[k]);
    }
    c->pool.free(p->d_order);
    c->pool.free(p->d_heads);
    c->pool.free(p->d_jobs);
    c->pool.free(p->d_rowmax);
    c->pool.free(p->d_plain_jobs);
    c->pool.free(p->d_plain_scratch);
    c->pool.free(p->d_corder);
    c->pool.free(p->d_declined);
    c->pool.free(p->d_plain_slot);
    delete p;
}

}  // namespace

int plan_batched(bwg_ctx *c, bwg_batch *b, hipStream_t st) { return plan_batch(c, b, st); }

int enqueue_batched(bwg_ctx *c, bwg_batch *b, StreamSet &ss, int64_t *launches) {
    *launches = 0;
    int rc = plan_batch(c, b, ss.main);
    if (rc) return rc;
    hipError_t e = hipSuccess;
    // Before the launch sequence is captured, learn from the eager run how many jobs each compacted
    // bin actually ran, so that the captured sequence deals the bins by their real work.
    if (b->tu.use_graphs && !b->tu.item_timing && !b->graph && b->batched_runs >= 1 && b->kind == 0) {
        e = learn_dp_shares(c, b, ss.main);
        if (e != hipSuccess) return bwg_map_hip(e);
    }
    // Capture the launch sequence once the batch has run eagerly; the eager run is also the
    // reference every later replay is compared with by the tests.
    if (b->tu.use_graphs && !b->tu.item_timing && !b->graph && b->batched_runs >= 1) {
        hipGraph_t g = nullptr;
        e = hipStreamBeginCapture(ss.main, hipStreamCaptureModeThreadLocal);
        int64_t nl = 0;
        const hipError_t e2 = e == hipSuccess ? emit_all(c, b, ss, &nl) : e;
        if (e == hipSuccess) e = hipStreamEndCapture(ss.main, &g);
        if (e == hipSuccess && e2 != hipSuccess) e = e2;
        if (e == hipSuccess) e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
        if (g) (void)hipGraphDestroy(g);
        if (e != hipSuccess) return bwg_map_hip(e);
    }
    e = hipEventRecord(b->ev_start, ss.main);
    if (e == hipSuccess) {
        if (b->graph && !b->tu.item_timing) {
            e = hipGraphLaunch(b->graph, ss.main);
            *launches = b->graph_launches;
        } else {
            e = emit_all(c, b, ss, launches);
            b->graph_launches = *launches;
        }
    }
    if (e == hipSuccess) e = hipEventRecord(b->ev_end, ss.main);
    b->batched_runs++;
    return bwg_map_hip(e);
}

size_t plan_device_bytes(bwg_ctx *c, int kind, size_t n, const uint32_t *q_len, const uint32_t *t_len,
                         const int32_t *h0, const ExtScoring *esc, const KswScoring *ksc, bool exact) {
    Tuning tu = c->tuning_snapshot();
    if (exact) tu.tune_new = false;
    auto W = [exact](size_t b) {
        return exact ? bump_round(std::max<size_t>(b, 1)) : DevicePool::worst(std::max<size_t>(b, 64));
    };
    // order array, the compacted queues of the ungapped pre-pass, queue heads and bin metadata
    size_t bytes = 2 * W(n * 4) + W(64 * (kCountStride + 3) * 4);
    if (kind == 0) {
        bytes += W(n);  // the ungapped pre-pass's verdicts
        size_t n_plain = 0, plain_cols = 0;
        std::vector<uint32_t> per_class(kExtShortClasses + kNumExtKs, 0);
        ExtClassifier classifier(*esc, tu);
        for (size_t k = 0; k < n; ++k) {
            ExtJob j{};
            j.qlen = q_len[k];
            j.tlen = t_len[k];
            j.h0 = h0[k];
            int ki = -1;
            const int cl = classifier.classify(j, &ki);
            if (cl == -1) {
                ++n_plain;
                plain_cols += (size_t)q_len[k] + 1;
            } else if (cl >= 0) {
                per_class[cl]++;
            }
        }
        if (n_plain) bytes += W(n_plain * sizeof(ExtJob)) + W(plain_cols * sizeof(int2));
        bool tunes = false;
        for (int cl = 0; cl < (int)per_class.size(); ++cl)
            tunes |= tu.autotune && tu.tune_new && per_class[cl] >= kTuneMinJobs &&
                     !c->tuner.lookup(scoring_sig(*esc), cl, nullptr);
        if (tunes) bytes += W((kTuneSample + 1) * 4);
        return bytes;
    }
    // rescue: the planned descriptor copy, the row-maximum tiles of the fast jobs (64 jobs a tile,
    // each tile as deep as its longest target: at most n_fast rounded up to 64 times the longest),
    // and the plain bin's per-job state
    const bool pk_scoring = ksw_pk_scoring_ok(*ksc);
    const int32_t mm = std::max<int32_t>(1, mat_max(ksc->mat));
    size_t n_fast = 0, n_plain = 0, plain_bytes = 0;
    uint64_t rows_fast = 0;
    uint32_t tmax_fast = 0;
    std::vector<uint32_t> members(16, 0), cls_tmax(16, 0);
    for (size_t k = 0; k < n; ++k) {
        const int qlen = (int)q_len[k];
        const int K = ksc->o_ins >= 0 && t_len[k] <= kMaxFastTlen ? ksw_k_for(widest_pad(qlen, ksc->max_sc)) : -1;
        if (K < 0) {
            ++n_plain;
            const size_t b = ((size_t)ksw_padded(qlen, ksc->max_sc) + 16) * 8 + (size_t)t_len[k] * 4;
            plain_bytes += (b + 15) & ~(size_t)15;
            continue;
        }
        ++n_fast;
        rows_fast += t_len[k];
        tmax_fast = std::max(tmax_fast, t_len[k]);
        const int cl = K / 2 - 1 + (pk_scoring && (int64_t)qlen * mm < kKswPkMaxScore ? 0 : 8);
        members[cl]++;
        cls_tmax[cl] = std::max(cls_tmax[cl], t_len[k]);
    }
    // Within a bin the jobs are sorted by descending target (by a key exact to 1/2048), so a tile
    // is no deeper than the shortest target of the tile before it: the tiles hold at most the jobs'
    // own rows (plus the key's rounding) and one tile of the longest target per bin (at most 16
    // fast bins) and one more.
    const uint64_t tiles = std::min<uint64_t>((n_fast + 63) / 64 * 64 * (uint64_t)tmax_fast,
                                              rows_fast + rows_fast / 1024 + 64ull * 17 * tmax_fast);
    bytes += W(n * sizeof(KswJob)) + W(tiles * 4);
    if (n_plain) bytes += W(n_plain * sizeof(KswJob)) + W(plain_bytes);
    size_t tune = 0;
    for (int cl = 0; cl < 16; ++cl) {
        if (!tu.autotune || !tu.tune_new || members[cl] < kTuneMinJobs ||
            c->tuner.lookup(scoring_sig(*ksc), cl, nullptr))
            continue;
        const size_t S = std::min<size_t>(members[cl], kTuneSample / 2);
        tune = std::max(tune, W((S + 1) * 4) + W(S * sizeof(KswJob)) + W((S + 63) / 64 * 64 * (size_t)cls_tmax[cl] * 4));
    }
    return bytes + tune;
}

const void *glb_plan_jobs(const bwg_batch *b) { return b->plan ? ((const Plan *)b->plan)->d_jobs : nullptr; }

int batched_ungapped_hits(bwg_ctx *c, bwg_batch *b, uint64_t *hits) {
    *hits = 0;
    if (!b->plan || b->kind != 0) return BWG_OK;
    Plan *p = (Plan *)b->plan;
    if (!p->ungapped) return BWG_OK;
    std::vector<uint32_t> cnt((size_t)kCountStride * p->n_bins);
    hipError_t e = hipMemcpyAsync(cnt.data(), p->d_heads + p->n_bins, cnt.size() * 4, hipMemcpyDeviceToHost,
                                  c->ss.main);
    if (e == hipSuccess) e = hipStreamSynchronize(c->ss.main);
    if (e != hipSuccess) return bwg_map_hip(e);
    uint64_t eligible = 0, declined = 0;
    for (const Item &it : p->items)
        if (is_ext_dp(it.kind)) {
            eligible += it.count;
            declined += cnt[(size_t)kCountStride * it.head];
        }
    *hits = eligible - declined;
    return BWG_OK;
}

std::string batched_report(bwg_batch *b) {
    std::string s;
    std::vector<Plan *> plans;
    if (b->plan) plans.push_back((Plan *)b->plan);
    for (void *q : b->retry_plans) plans.push_back((Plan *)q);
    char line[256], name[64];
    int idx = 0;
    for (size_t pi = 0; pi < plans.size(); ++pi) {
        Plan *p = plans[pi];
        for (size_t k = 0; k < p->items.size(); ++k, ++idx) {
            const Item &it = p->items[k];
            double us = -1.0;
            if (k < p->ev0.size() && p->ev0[k] && p->ev1[k]) {
                float ms = 0.f;
                if (hipEventElapsedTime(&ms, p->ev0[k], p->ev1[k]) == hipSuccess) us = ms * 1e3;
            }
            const int width = (it.kind == kKswFwd || it.kind == kKswRev || it.kind == kGlbGroup ||
                               is_ext_dp(it.kind))
                                  ? (it.narrow ? 16 : 32)
                                  : 32;
            std::snprintf(line, sizeof line, "%d round=%d stage=%d %s jobs=%u grid=%d us=%.3f\n", idx,
                          p->retry_round, it.stage, mapping_name(it.kind, it.param, width, name, sizeof name),
                          it.count, it.launch_grid > 0 ? it.launch_grid : it.grid, us);
            s += line;
        }
    }
    return s;
}

int batched_job_bins(bwg_batch *b, int32_t *bin) {
    for (size_t k = 0; k < b->n; ++k) bin[k] = -1;
    Plan *p = (Plan *)b->plan;
    if (!p) return BWG_EINVAL;
    for (size_t r = 0; r < p->items.size(); ++r) {
        const Item &it = p->items[r];
        const bool is_bin = is_ext_dp(it.kind) || it.kind == kExtPlain ||
                            it.kind == kKswFwd || it.kind == kKswPlain || it.kind == kGlbGroup ||
                            it.kind == kGlbPlain;
        if (!is_bin) continue;
        for (uint32_t q = 0; q < it.count; ++q) bin[p->order[it.offset + q]] = (int32_t)r;
    }
    return BWG_OK;
}

void release_batched(bwg_ctx *c, bwg_batch *b) {
    if (b->graph) {
        (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    for (void *q : b->retry_plans) free_plan(c, (Plan *)q);
    b->retry_plans.clear();
    if (!b->plan) return;
    free_plan(c, (Plan *)b->plan);
    b->plan = nullptr;
}

}  // namespace bwg
