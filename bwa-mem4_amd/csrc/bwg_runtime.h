// Host-side runtime structures: the context, device memory pool and staged batches.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <map>
#include <mutex>
#include <vector>

#include "bwg_internal.h"

namespace bwg {

constexpr uint32_t kMaxQlen = 4096;
constexpr uint32_t kMaxTlen = 65535;

// Caching device allocator: freed blocks are kept per rounded size and reused, so a steady
// stream of batches does not call hipMalloc/hipFree on the hot path.
class DevicePool {
  public:
    void *alloc(size_t bytes);
    void free(void *p);
    void release_all();
    size_t bytes_in_use() const { return in_use_; }
    size_t bytes_cached() const { return cached_; }

  private:
    static size_t round(size_t b);
    std::mutex mu_;
    std::multimap<size_t, void *> free_;
    std::map<void *, size_t> live_;
    size_t in_use_ = 0, cached_ = 0;
};

}  // namespace bwg

struct bwg_ctx {
    int device = 0;
    int num_cus = 256;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    std::mutex run_mu;  // one batch on the context stream at a time
    int form = BWG_FORM_BATCHED;
    std::atomic<int64_t> launches{0}, jobs{0};
    int64_t last_kernel_ns = 0, last_main_ns = 0, last_main_launches = 0;
    bwg::DevicePool pool;
};

struct bwg_batch {
    int kind = 0;  // 0 extension, 1 rescue
    size_t n = 0;
    bwg::ExtScoring esc{};
    bwg::KswScoring ksc{};
    uint8_t *d_arena = nullptr;
    size_t arena_len = 0;
    void *d_jobs = nullptr;  // ExtJob[n] or KswJob[n] in arrival order
    void *d_out = nullptr;   // bwg_ext_result[n] or bwg_ksw_result[n]
    // host copies used to plan the launches
    std::vector<uint32_t> q_len, t_len;
    std::vector<int32_t> h0;
    // per-form device state, created on first run in that form
    void *d_plain_scratch = nullptr;
    size_t plain_scratch_bytes = 0;
    void *d_plain_jobs = nullptr;  // descriptors carrying the plain form's scratch offsets
    void *ext_plan = nullptr;      // batched form: bins, order and queues (bwg_batched.cpp)
    void *ksw_plan = nullptr;
};

namespace bwg {
int stage_ext(bwg_ctx *c, const ExtScoring &sc, size_t n, const uint8_t *arena, size_t arena_len,
              const uint64_t *q_off, const uint32_t *q_len, const uint64_t *t_off,
              const uint32_t *t_len, const int32_t *h0, bwg_batch **out);
int stage_ksw(bwg_ctx *c, const KswScoring &sc, size_t n, const uint8_t *arena, size_t arena_len,
              const uint64_t *q_off, const uint32_t *q_len, const uint64_t *t_off,
              const uint32_t *t_len, bwg_batch **out);
int run_batch(bwg_ctx *c, bwg_batch *b);
int fetch_results(bwg_ctx *c, bwg_batch *b, void *out);
void destroy_batch(bwg_ctx *c, bwg_batch *b);
}  // namespace bwg

int bwg_map_hip(hipError_t e);
