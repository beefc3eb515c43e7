// Batched form: device-side length binning, per-bin kernel mapping chosen by the autotuner,
// persistent work-queue kernels.
#pragma once
#include "bwg_runtime.h"

namespace bwg {

struct BatchedTiming {
    int64_t total_ns = 0, main_ns = 0, launches = 0, main_launches = 0;
};

int run_batched(bwg_ctx *c, bwg_batch *b, BatchedTiming *t);
void release_batched(bwg_ctx *c, bwg_batch *b);
int run_batched_ksw(bwg_ctx *c, bwg_batch *b, BatchedTiming *t);
void release_batched_ksw(bwg_ctx *c, bwg_batch *b);

}  // namespace bwg
