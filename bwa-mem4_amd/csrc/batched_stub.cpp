#include "bwg_batched.h"
namespace bwg {
int run_batched(bwg_ctx *, bwg_batch *, BatchedTiming *) { return BWG_EINVAL; }
void release_batched(bwg_ctx *, bwg_batch *) {}
}  // namespace bwg
