// Host-side runtime structures: the context, device memory pool, stream sets and staged batches.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <map>
#include <mutex>
#include <vector>

#include "bwg_batched_kernels.h"
#include "bwg_internal.h"
#include "bwg_tuner.h"

namespace bwg {

constexpr uint32_t kMaxQlen = 4096;
constexpr uint32_t kMaxTlen = 65535;
constexpr int kAuxStreams = 3;

// Caching device allocator: freed blocks are kept per rounded size and reused, so a steady
// stream of batches does not call hipMalloc/hipFree on the hot path.
class DevicePool {
  public:
    void *alloc(size_t bytes);
    void free(void *p);
    void release_all();
    size_t bytes_in_use() const { return in_use_; }
    size_t bytes_cached() const { return cached_; }
    static size_t round(size_t b);

  private:
    std::mutex mu_;
    std::multimap<size_t, void *> free_;
    std::map<void *, size_t> live_;
    size_t in_use_ = 0, cached_ = 0;
};

// A main stream plus helper streams used to run independent kernels of one batch concurrently
// (fork after a main-stream event, join back before the batch's end event).
struct StreamSet {
    hipStream_t main = nullptr;
    hipStream_t aux[kAuxStreams] = {nullptr, nullptr, nullptr};
    hipEvent_t fork = nullptr;
    hipEvent_t join[kAuxStreams] = {nullptr, nullptr, nullptr};
    int create();
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
};

}  // namespace bwg

struct bwg_ctx {
    int device = 0;
    int num_cus = 256;
    bwg::StreamSet ss;           // the context's own streams (synchronous entry points)
    std::mutex run_mu;           // one synchronous batch on the context streams at a time
    int form = BWG_FORM_BATCHED;
    bwg::Tuning tuning;
    bwg::Tuner tuner;
    std::atomic<int64_t> launches{0}, jobs{0};
    int64_t last_kernel_ns = 0, last_main_ns = 0, last_main_launches = 0;
    bwg::DevicePool pool;
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
    bwg::ExtScoring esc{};
    bwg::KswScoring ksc{};
    // host copies of the job descriptors, used to plan the launches
    std::vector<bwg::ExtJob> hext;
    std::vector<bwg::KswJob> hksw;
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
};

namespace bwg {

// Stage a batch: copy arena and descriptors to the device (asynchronously on `st`).  `ref` (may be
// null) is a device reference array for jobs whose targets are reference windows (bit 62 of
// t_off; bit 63 reads the window back to front).
int stage_ext(bwg_ctx *c, const ExtScoring &sc, size_t n, const uint8_t *arena, size_t arena_len,
              const uint64_t *q_off, const uint32_t *q_len, const uint64_t *t_off,
              const uint32_t *t_len, const int32_t *h0, const uint8_t *ref, hipStream_t st,
              bwg_batch **out);
// A one-sided extension batch with the reference's band doubling: up to `tries` rounds at bands
// sc.w << r; prev holds each job's score before the first round.
int stage_side(bwg_ctx *c, const ExtScoring &sc, int tries, size_t n, const uint8_t *arena, size_t arena_len,
               const uint64_t *q_off, const uint32_t *q_len, const uint64_t *t_off, const uint32_t *t_len,
               const int32_t *h0, const int32_t *prev, hipStream_t st, bwg_batch **out);
int fetch_side_bands(bwg_ctx *c, bwg_batch *b, int32_t *w_out);
// A mate-rescue round: n anchors, 4n jobs whose windows the device computes against the context's
// two-strand reference.
int stage_matesw(bwg_ctx *c, const KswScoring &sc, const MateRescueParams &mp, size_t n,
                 const bwg_mate_anchor *anchors, const int64_t *ma_rb, size_t n_ma, const uint8_t *arena,
                 size_t arena_len, hipStream_t st, bwg_batch **out);
int fetch_windows(bwg_ctx *c, bwg_batch *b, bwg_rescue_window *win);
// Banded global alignment with traceback: n jobs with their own bands.
int stage_global(bwg_ctx *c, const GlobalScoring &sc, size_t n, const uint8_t *arena, size_t arena_len,
                 const uint64_t *q_off, const uint32_t *q_len, const uint64_t *t_off, const uint32_t *t_len,
                 const int32_t *w, hipStream_t st, bwg_batch **out);
int fetch_cigar_lengths(bwg_ctx *c, bwg_batch *b, uint32_t *n_ops);
int fetch_cigars(bwg_ctx *c, bwg_batch *b, uint32_t *ops, size_t cap);
int stage_ksw(bwg_ctx *c, const KswScoring &sc, size_t n, const uint8_t *arena, size_t arena_len,
              const uint64_t *q_off, const uint32_t *q_len, const uint64_t *t_off,
              const uint32_t *t_len, const uint8_t *ref, hipStream_t st, bwg_batch **out);
// Enqueue the batch's kernels on the stream set (no host synchronisation).  b->ev_start/ev_end
// bracket them on ss.main.
int enqueue_batch(bwg_ctx *c, bwg_batch *b, int form, StreamSet &ss);
// Synchronous run on the context streams: enqueue, wait, update the context statistics.
int run_batch(bwg_ctx *c, bwg_batch *b);
int enqueue_results(bwg_ctx *c, bwg_batch *b, void *out, hipStream_t st);
int fetch_results(bwg_ctx *c, bwg_batch *b, void *out);
void destroy_batch(bwg_ctx *c, bwg_batch *b);
// Device bytes a batch of this description holds once staged and planned (admission control).
size_t batch_device_bytes(int kind, size_t n, size_t arena_len, const uint32_t *q_len,
                          const uint32_t *t_len, int32_t max_sc);

// batched form (bwg_batched.cpp)
int enqueue_batched(bwg_ctx *c, bwg_batch *b, StreamSet &ss, int64_t *launches);
// The planned descriptors of a global-alignment batch (device), null before its first run.
const void *glb_plan_jobs(const bwg_batch *b);
void release_batched(bwg_ctx *c, bwg_batch *b);
// Jobs of the last batched run that the ungapped fast path resolved (0 when it did not run).
int batched_ungapped_hits(bwg_ctx *c, bwg_batch *b, uint64_t *hits);

}  // namespace bwg

int bwg_map_hip(hipError_t e);
