// Target-code prefetch for the one-job-per-lane kernel.  A DP row needs one target code, and a short
// row takes less time than a global load, so fetching the next row's code one row ahead leaves every
// row waiting on memory.  The stream fetches codes 8 rows ahead into registers that are only read
// once the block they belong to starts, so the load latency is hidden behind the rows in between.
// Indices past the end of the target are clamped to its last code (loads stay in bounds); those
// codes are never used.
#pragma once
#include <hip/hip_runtime.h>

#include "bwg_internal.h"

namespace bwg {

// One stream per lane (one job per lane): the codes of rows [r0, r0+8) packed in lo/hi, the codes
// of rows [r0+8, r0+16) in flight.  Rows must be requested in order 0, 1, 2, ...
struct LaneTargetStream {
    uint32_t lo, hi;
    uint32_t n[8];
    __device__ __forceinline__ void start(const TargetView &t, int32_t tlen) {
        uint32_t c[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) c[k] = t[min(k, tlen - 1)];
        lo = c[0] | c[1] << 8 | c[2] << 16 | c[3] << 24;
        hi = c[4] | c[5] << 8 | c[6] << 16 | c[7] << 24;
#pragma unroll
        for (int k = 0; k < 8; ++k) n[k] = t[min(8 + k, tlen - 1)];
    }
    __device__ __forceinline__ uint32_t at(int32_t r, const TargetView &t, int32_t tlen) {
        if ((r & 7) == 0 && r > 0) {
            lo = n[0] | n[1] << 8 | n[2] << 16 | n[3] << 24;
            hi = n[4] | n[5] << 8 | n[6] << 16 | n[7] << 24;
#pragma unroll
            for (int k = 0; k < 8; ++k) n[k] = t[min(r + 8 + k, tlen - 1)];
        }
        return __builtin_amdgcn_perm(hi, lo, (uint32_t)(r & 7)) & 0xffu;
    }
};

}  // namespace bwg
