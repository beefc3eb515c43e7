// Host-side runtime structures: the context, device memory pool, stream sets and staged batches.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "bwg_batched_kernels.h"
#include "bwg_internal.h"
#include "bwg_tuner.h"

namespace bwg {

// Job sizes.  Every job up to kMaxJobLen bases of query and of target runs on the device: the
// binned kernels take the shapes they are tuned for (extension queries below 8192 columns with
// h0 + qlen*max_sc < 2^18; rescue queries padded to at most 256 columns and targets up to
// kMaxFastTlen rows), and the plain kernels, whose per-job state lives in global memory, take
// everything else.  The limit keeps every score, row index and gap product of the int32 recurrences
// away from overflow (checked per job on the host, see bwg_api.cpp).
constexpr uint32_t kMaxJobLen = 1u << 24;
constexpr uint32_t kMaxFastTlen = 65535;
constexpr int kAuxStreams = 3;

// Device scratch of the allocation-free entry point: one array per GPU for the life of the process
// (noalloc.hip).  A call that needs more than this is rejected with BWG_E2BIG.
constexpr size_t kNoAllocScratchBytes = 64ull << 20;
constexpr size_t kBumpAlign = 256;
inline size_t bump_round(size_t b) { return (b + kBumpAlign - 1) & ~(kBumpAlign - 1); }

// Bump allocation over a fixed device region (the no-alloc scratch): nothing is ever freed during
// a call; the region is reused by the next call.
struct Bump {
    uint8_t *base = nullptr;
    size_t cap = 0, used = 0;
    void *take(size_t bytes) {
        const size_t b = bump_round(std::max<size_t>(bytes, 1));
        if (b > cap - used) return nullptr;
        void *p = base + used;
        used += b;
        return p;
    }
};

// Size limits of one job (BWG_E2BIG beyond them): both lengths up to kMaxJobLen, and the largest
// score and gap product the int32 recurrences can form kept below 2^30.
inline bool ext_job_fits(const ExtScoring &sc, uint32_t qlen, uint32_t tlen, int32_t h0) {
    if (qlen > kMaxJobLen || tlen > kMaxJobLen) return false;
    const int64_t top = (int64_t)h0 + (int64_t)qlen * std::max(sc.max_sc, 0) + std::max(sc.end_bonus, 0);
    const int64_t gap = (int64_t)std::max(sc.e_del, sc.e_ins) * ((int64_t)std::max(qlen, tlen) + 1) +
                        std::max(std::max(sc.o_del, sc.o_ins), 0);
    return top < (1 << 30) && gap < (1 << 30);
}
inline bool ksw_job_fits(const KswScoring &sc, uint32_t qlen, uint32_t tlen) {
    if (qlen > kMaxJobLen || tlen > kMaxJobLen) return false;
    const int64_t top = ((int64_t)qlen + 16) * std::max<int32_t>(mat_max(sc.mat), 0);
    const int64_t gap = (int64_t)std::max(sc.e_del, sc.e_ins) * ((int64_t)std::max(qlen, tlen) + 16) +
                        std::max(std::max(sc.o_del, sc.o_ins), 0);
    return top < (1 << 30) && gap < (1 << 30);
}

// Caching device allocator: freed blocks are kept per rounded size and reused, so a steady
// stream of batches does not call hipMalloc/hipFree on the hot path.  A cached block is reused for
// a request whose rounded size it exceeds by at most a quarter (kReuseSlack), and the cache is
// trimmed, largest blocks first, whenever it would hold more than its cap.
class DevicePool {
  public:
    void *alloc(size_t bytes);
    void free(void *p);
    void release_all();
    size_t bytes_in_use() const { return in_use_.load(); }
    size_t bytes_cached() const { return cached_.load(); }
    size_t peak_in_use() const { return peak_.load(); }
    void reset_peak() { peak_.store(in_use_.load()); }
    void set_cache_cap(size_t cap);
    size_t cache_cap() const { return cap_.load(); }
    static size_t round(size_t b);
    // Device bytes one alloc(bytes) can hold, including the reuse slack: the admission estimate.
    static size_t worst(size_t b) { return round(b) + (round(b) >> 2); }

  private:
    void trim_locked(size_t cap);
    std::mutex mu_;
    std::multimap<size_t, void *> free_;
    std::map<void *, size_t> live_;
    std::atomic<size_t> in_use_{0}, cached_{0}, peak_{0}, cap_{size_t(8) << 30};
};

// A main stream plus helper streams used to run independent kernels of one batch concurrently
// (fork after a main-stream event, join back before the batch's end event).  The runtime maps
// streams onto a few hardware queues, and a set's streams created in a row take consecutive
// queues; so that the main streams of several sets (execution lanes, server slots) do not all
// share one queue, and so run one after another, set k takes the (k mod 4)-th of its four streams
// as its main one.
struct StreamSet {
    hipStream_t main = nullptr;
    hipStream_t aux[kAuxStreams] = {nullptr, nullptr, nullptr};
    hipEvent_t fork = nullptr;
    hipEvent_t join[kAuxStreams] = {nullptr, nullptr, nullptr};
    int create(int rotation = 0);
    void destroy();
};

struct Tuning {
    int ext_thread_max_q16 = 16;   // queries up to this length run one job per lane, 16-column LDS
    int ext_thread_max_q32 = 32;   // ... 32-column LDS
    int ext_group_max_qcap = 512;  // longest query the 16-lane group kernel keeps in LDS
    int ext_group_oversub = 1;     // persistent grid = oversub * occupancy * CUs
    int ksw_group_oversub = 1;
    bool autotune = true;          // time candidate mappings per length class (bwg_tuner.h)
    bool ext_k_step = false;       // untuned extension bins: one K above the minimal one
    bool ksw_k_step = false;       // untuned rescue bins: one K above the minimal one
    bool ext_force_wide = false;   // extension: 32-bit column state even where 16 bits are exact
    bool ksw_force_wide = false;   // rescue: 32-bit DP state even where 16 bits are exact
    bool ungapped = true;          // extension: the reference's ungapped-diagonal fast path first
    bool concurrent_bins = true;   // run the bins of one batch on the stream set concurrently
    bool use_graphs = true;        // replay a batch's launch sequence as a captured hipGraph
    bool tune_new = true;          // a batch may run the autotuner (false: it only reads the table)
    bool item_timing = false;      // batched runs time each launch with events (no graph replay)
    bool coalesce = true;          // concurrent small synchronous calls are merged (bwg_api.cpp)
    int host_threads = 0;          // BWG_OPT_HOST_THREADS: 0 = the CPUs this process may use, at most 16
    int ext_group_lanes = 16;      // BWG_OPT_EXT_GROUP_LANES: lanes per job of untuned group bins (16 or 8)
    int ksw_group_lanes = 16;      // BWG_OPT_KSW_GROUP_LANES: lanes per job of untuned packed rescue bins
    int bin_streams = kAuxStreams + 1;  // BWG_OPT_BIN_STREAMS: streams a stage's concurrent bins are dealt onto
    int form = BWG_FORM_BATCHED;   // BWG_OPT_FORM
};

// Host passes over the jobs of a large call (validation, descriptors, classification, binning)
// split into contiguous ranges over worker threads: parts(tu, n) ranges of at least kHostGrain
// jobs, fn(begin, end, part) on part-1 new threads and the caller.
constexpr size_t kHostGrain = 65536;

// An allocator whose resize leaves new elements default-initialised (no zeroing pass): the host
// arrays of a large call are written in full, in parallel, right after they are sized.
template <class T>
struct NoInitAlloc : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = NoInitAlloc<U>;
    };
    NoInitAlloc() = default;
    template <class U>
    NoInitAlloc(const NoInitAlloc<U> &) {}
    template <class U>
    void construct(U *p) {
        ::new ((void *)p) U;
    }
    template <class U, class... A>
    void construct(U *p, A &&...a) {
        ::new ((void *)p) U(std::forward<A>(a)...);
    }
};
template <class T>
using HostVec = std::vector<T, NoInitAlloc<T>>;
int host_parts(const Tuning &tu, size_t n);
// fn(part) for every part in [0, parts), on the process's host worker threads and the caller (which
// takes parts too, so a call completes even when every worker is busy); returns when all are done.
void run_parts(int parts, const std::function<void(int)> &fn);
template <class Fn>
void parallel_for(size_t n, int parts, Fn &&fn) {
    if (parts <= 1) {
        fn((size_t)0, n, 0);
        return;
    }
    run_parts(parts, [&fn, n, parts](int q) { fn(n * q / parts, n * (q + 1) / parts, q); });
}

// An execution lane of the synchronous entry points: its own stream set, so that calls from
// several host threads stage, run and copy back concurrently instead of queueing behind one
// another, plus pinned result staging.
struct ExecLane {
    StreamSet ss;
    void *pin_out = nullptr;
    size_t pin_out_cap = 0;
    int create(int index);
    void destroy();
};

// Pinned chunks through which a large pageable arena goes to the device (upload_pinned): two
// buffers for each of up to four copying threads.  A context keeps a few such sets (acquire_upload)
// and reuses them; a buffer is refilled only after its last DMA has completed.
struct UploadStaging {
    static constexpr int kThreads = 4;
    static constexpr int kBufs = 2 * kThreads;
    static constexpr size_t kChunk = 8u << 20;
    void *buf[kBufs] = {};
    hipEvent_t ev[kBufs] = {};
    bool pending[kBufs] = {};
    int create();
    void destroy();
};

// Host time of the phases of the last synchronous call (bwg_extend_batch, bwg_ksw_align2_batch).
enum Phase { kPhValidate, kPhStage, kPhPlan, kPhKernels, kPhFetch, kPhTotal, kNumPhases };

}  // namespace bwg

struct bwg_ctx {
    int device = 0;
    int num_cus = 256;
    bwg::StreamSet ss;           // the context's own streams (staged batches)
    std::mutex run_mu;           // one staged-batch operation on the context streams at a time
    // options: written by bwg_set_option under opt_mu; every batch plans with a copy taken under
    // the same lock (tuning_snapshot), so a concurrent option change never tears a plan
    mutable std::mutex opt_mu;
    bwg::Tuning tuning;
    bwg::Tuning tuning_snapshot() const {
        std::lock_guard<std::mutex> g(opt_mu);
        return tuning;
    }
    bwg::Tuner tuner;
    // share of a bin's jobs that reach the DP after the extension's ungapped pre-pass, by bin code
    // (kind << 8 | param << 1 | 16-bit state), measured on earlier runs; < 0 when not yet seen.  It
    // weights the bins' nominal work when a launch sequence deals them onto the streams.
    static constexpr int kDpShareCodes = 16 << 8;
    std::mutex dp_mu;
    float dp_share[kDpShareCodes];
    std::atomic<int64_t> launches{0}, jobs{0};
    std::atomic<int64_t> last_kernel_ns{0}, last_main_ns{0}, last_main_launches{0};
    std::atomic<int64_t> phase_ns[bwg::kNumPhases] = {};
    bwg::DevicePool pool;
    // execution lanes of the synchronous entry points (created on demand, at most max_lanes)
    std::mutex lane_mu;
    std::condition_variable lane_cv;
    std::vector<std::unique_ptr<bwg::ExecLane>> lanes;
    std::vector<bwg::ExecLane *> free_lanes;
    int max_lanes = 4;
    // timing events of the allocation-free entry point, created with the context
    hipEvent_t na_ev[2] = {nullptr, nullptr};
    // pinned upload staging of the synchronous calls (at most kMaxUploadSets, made on demand)
    static constexpr int kMaxUploadSets = 4;
    std::mutex up_mu;
    std::vector<std::unique_ptr<bwg::UploadStaging>> up_all;
    std::vector<bwg::UploadStaging *> up_free;
    // synchronous calls in progress, and the server that merges concurrent small ones
    std::atomic<int> sync_active{0};
    std::mutex coalesce_mu;
    bwg_server *coalescer = nullptr;
    // two-strand reference for mate rescue (bwg_set_reference): 2*l_pac bytes and the contigs
    uint8_t *d_ref2 = nullptr;
    int64_t l_pac = 0;
    int64_t *d_contig_off = nullptr;
    int32_t *d_contig_len = nullptr;
    int32_t n_contigs = 0;
    std::atomic<int> ref_users{0};  // staged rescue batches reading the reference
};

struct bwg_batch {
    int kind = 0;  // 0 extension, 1 rescue, 2 global alignment with traceback
    size_t n = 0;
    // a large arena's upload, issued from a helper thread while the calling thread plans the batch
    // (synchronous entry points only); wait_arena joins it before anything reads the arena
    std::thread arena_copy;
    hipError_t arena_err = hipSuccess;
    bwg::ExtScoring esc{};
    bwg::KswScoring ksc{};
    // host copies of the job descriptors, used to plan the launches
    bwg::HostVec<bwg::ExtJob> hext;
    bwg::HostVec<bwg::KswJob> hksw;
    std::vector<bwg::GlobalJob> hglb;
    bwg::GlobalScoring gsc{};
    uint32_t *d_z = nullptr;      // global: traceback codes
    uint32_t *d_ops = nullptr;    // global: CIGAR runs back to front, per-job scratch
    uint32_t *d_nops = nullptr;   // global: runs per job
    // device buffers
    uint8_t *d_arena = nullptr;  // query bytes, and target bytes unless targets come from a reference
    size_t arena_len = 0;
    const uint8_t *d_ref = nullptr;  // device-resident reference for targets given as windows
    void *d_jobs = nullptr;
    void *d_out = nullptr;
    // plain form state, created on first plain run
    void *d_plain_scratch = nullptr;
    void *d_plain_jobs = nullptr;
    // band-retry extension (bwg_batch_stage_side): band tries, the callers' previous scores, and
    // per-job device state between rounds
    int side_tries = 0;
    int32_t *d_prev0 = nullptr, *d_prev = nullptr, *d_wout = nullptr;
    uint8_t *d_active = nullptr;
    // mate rescue with device-resident windows (bwg_batch_stage_matesw): 4 jobs per anchor
    bool matesw = false;
    uint32_t n_anchors = 0;
    bwg::MateRescueParams mp{};
    bwg::MateAnchor *d_anchors = nullptr;
    int64_t *d_ma_rb = nullptr;
    bwg_rescue_window *d_win = nullptr;
    const int64_t *d_contig_off = nullptr;
    const int32_t *d_contig_len = nullptr;
    bwg_ctx *ref_owner = nullptr;  // the context whose reference this batch holds a use of
    // batched form state (bwg_batched.cpp)
    void *plan = nullptr;
    std::vector<void *> retry_plans;  // band-retry rounds 1.. of a side batch
    hipGraphExec_t graph = nullptr;
    int64_t graph_launches = 0;
    int batched_runs = 0;
    // timing of the last run
    hipEvent_t ev_start = nullptr, ev_end = nullptr;
    int64_t last_launches = 0;
    // the options the batch was planned with (a snapshot taken when it was staged)
    bwg::Tuning tu;
    // allocation-free batches: device buffers come from this region instead of the pool, and the
    // timing events belong to the context
    bwg::Bump *bump = nullptr;
    bool borrowed_events = false;
};

namespace bwg {

// Stage a batch: copy arena and descriptors to the device (asynchronously on `st`).  `ref` (may be
// null) is a device reference array for jobs whose targets are reference windows (bit 62 of
// t_off; bit 63 reads the window back to front).
int stage_ext(bwg_ctx *c, const ExtScoring &sc, size_t n, const uint8_t *arena, size_t arena_len,
              const uint64_t *q_off, const uint32_t *q_len, const uint64_t *t_off,
              const uint32_t *t_len, const int32_t *h0, const uint8_t *ref, hipStream_t st,
              bwg_batch **out, bool defer_arena = false);
// A one-sided extension batch with the reference's band doubling: up to `tries` rounds at bands
// sc.w << r; prev holds each job's score before the first round.
int stage_side(bwg_ctx *c, const ExtScoring &sc, int tries, size_t n, const uint8_t *arena, size_t arena_len,
               const uint64_t *q_off, const uint32_t *q_len, const uint64_t *t_off, const uint32_t *t_len,
               const int32_t *h0, const int32_t *prev, hipStream_t st, bwg_batch **out, bool defer_arena = false);
int fetch_side_bands(bwg_ctx *c, bwg_batch *b, int32_t *w_out, hipStream_t st);
// A mate-rescue round: n anchors, 4n jobs whose windows the device computes against the context's
// two-strand reference.
int stage_matesw(bwg_ctx *c, const KswScoring &sc, const MateRescueParams &mp, size_t n,
                 const bwg_mate_anchor *anchors, const int64_t *ma_rb, size_t n_ma, const uint8_t *arena,
                 size_t arena_len, hipStream_t st, bwg_batch **out);
int fetch_windows(bwg_ctx *c, bwg_batch *b, bwg_rescue_window *win, hipStream_t st);
// Banded global alignment with traceback: n jobs with their own bands.
int stage_global(bwg_ctx *c, const GlobalScoring &sc, size_t n, const uint8_t *arena, size_t arena_len,
                 const uint64_t *q_off, const uint32_t *q_len, const uint64_t *t_off, const uint32_t *t_len,
                 const int32_t *w, hipStream_t st, bwg_batch **out);
// The fetch_* calls below enqueue their copies (and the CIGAR gather) on `st` and wait for them.
int fetch_cigar_lengths(bwg_ctx *c, bwg_batch *b, uint32_t *n_ops, hipStream_t st);
int fetch_cigars(bwg_ctx *c, bwg_batch *b, uint32_t *ops, size_t cap, hipStream_t st);
int stage_ksw(bwg_ctx *c, const KswScoring &sc, size_t n, const uint8_t *arena, size_t arena_len,
              const uint64_t *q_off, const uint32_t *q_len, const uint64_t *t_off,
              const uint32_t *t_len, const uint8_t *ref, hipStream_t st, bwg_batch **out, bool defer_arena = false);
// Host-side preparation of a batch for a form: the batched form's plan (binning, autotuning, plan
// uploads), or the plain form's per-job scratch.  Done once; enqueue_batch does it if needed.
int prepare_batch(bwg_ctx *c, bwg_batch *b, int form, StreamSet &ss);
// Enqueue the batch's kernels on the stream set (no host synchronisation).  b->ev_start/ev_end
// bracket them on ss.main.
int enqueue_batch(bwg_ctx *c, bwg_batch *b, int form, StreamSet &ss);
// Synchronous run on the context streams: enqueue, wait, update the context statistics.
int run_batch(bwg_ctx *c, bwg_batch *b);
// After a batch enqueued on some stream set has completed: record its kernel time and launches in
// the context statistics.
void record_run(bwg_ctx *c, bwg_batch *b);
int enqueue_results(bwg_ctx *c, bwg_batch *b, void *out, hipStream_t st);
int fetch_results(bwg_ctx *c, bwg_batch *b, void *out);
void destroy_batch(bwg_ctx *c, bwg_batch *b);
// Execution lanes of the synchronous entry points: take one (creating it, or waiting for one to
// free up), give it back.
ExecLane *acquire_lane(bwg_ctx *c);
void release_lane(bwg_ctx *c, ExecLane *l);
// Device memory of a batch: from its bump region when it has one, else from the context's pool.
void *batch_alloc(bwg_ctx *c, bwg_batch *b, size_t bytes);
// Joins a deferred arena upload (a no-op otherwise) and returns its status.
int wait_arena(bwg_batch *b);
// A free pinned upload set of the context (nullptr when all are in use), and its return.
UploadStaging *acquire_upload(bwg_ctx *c);
void release_upload(bwg_ctx *c, UploadStaging *u);
void destroy_uploads(bwg_ctx *c);
// len bytes of pageable host memory to dst through u's pinned chunks, on up to four threads, each
// chunk's DMA enqueued on st; returns once every chunk is enqueued (they complete in stream order).
hipError_t upload_pinned(UploadStaging &u, uint8_t *dst, const uint8_t *src, size_t len, hipStream_t st,
                         int device);
// The allocation-free entry point's scratch (noalloc.hip): its device address is looked up once
// per device (noalloc_setup, with that device current); calls on one device are serialised by
// noalloc_call_mutex, since they share the array.
int noalloc_setup(int device);
uint8_t *noalloc_scratch(int device);
std::mutex &noalloc_call_mutex(int device);
// Stage an extension batch whose buffers come from `bump` and whose events are `ev`.
int stage_ext_bump(bwg_ctx *c, const ExtScoring &sc, size_t n, const uint8_t *arena, size_t arena_len,
                   const uint64_t *q_off, const uint32_t *q_len, const uint64_t *t_off, const uint32_t *t_len,
                   const int32_t *h0, Bump *bump, const hipEvent_t ev[2], hipStream_t st, bwg_batch **out);
// Upper bound of the device bytes a batch of this description holds once staged and planned,
// tuning included (admission control of the server).  Mirrors the allocations of stage_ext /
// stage_ksw and of the planner (bwg_batched.cpp).
size_t batch_device_bytes(bwg_ctx *c, int kind, size_t n, size_t arena_len, const uint32_t *q_len,
                          const uint32_t *t_len, const int32_t *h0, const ExtScoring *esc,
                          const KswScoring *ksc);

// batched form (bwg_batched.cpp)
int enqueue_batched(bwg_ctx *c, bwg_batch *b, StreamSet &ss, int64_t *launches);
int plan_batched(bwg_ctx *c, bwg_batch *b, hipStream_t st);
// The planner's share of batch_device_bytes: order arrays, queues, plain-bin scratch, rescue row
// maxima, and the autotuner's sample buffers when the batch would tune a class.
// With `exact` set (allocation-free extension batches, which never tune): the bytes the plan takes
// from a bump region.
size_t plan_device_bytes(bwg_ctx *c, int kind, size_t n, const uint32_t *q_len, const uint32_t *t_len,
                         const int32_t *h0, const ExtScoring *esc, const KswScoring *ksc, bool exact = false);
// The planned descriptors of a global-alignment batch (device), null before its first run.
const void *glb_plan_jobs(const bwg_batch *b);
void release_batched(bwg_ctx *c, bwg_batch *b);
// Jobs of the last batched run that the ungapped fast path resolved (0 when it did not run).
int batched_ungapped_hits(bwg_ctx *c, bwg_batch *b, uint64_t *hits);
// The launches of a batch's last batched run, one line each (bwg_batch_report), and each job's
// launch index (bwg_batch_job_bins).
std::string batched_report(bwg_batch *b);
int batched_job_bins(bwg_batch *b, int32_t *bin);

}  // namespace bwg

int bwg_map_hip(hipError_t e);
