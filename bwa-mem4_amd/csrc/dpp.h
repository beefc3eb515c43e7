// Cross-lane primitives on 16-lane DPP rows (gfx9 DPP controls).  A "group" is one 16-lane DPP
// row of a 64-wide wavefront; all reductions and scans here stay inside the row.
#pragma once
#include <hip/hip_runtime.h>

namespace bwg {

// DPP control codes (gfx9 encoding).
enum : int {
    kRowShr = 0x110,  // + n : lane l reads lane l-n of its row
    kRowRor = 0x120,  // + n : lane l reads lane (l-n) mod 16 of its row
};

template <int CTRL, bool BOUND_ZERO>
__device__ __forceinline__ int dpp(int old, int src) {
    return __builtin_amdgcn_update_dpp(old, src, CTRL, 0xf, 0xf, BOUND_ZERO);
}

// Max over the 16 lanes of the row, result in every lane of the row.  A rotation reads no lane
// outside the row, so the bound-control fill never applies; asking for it lets the compiler fold
// each move into the max as a DPP operand.
__device__ __forceinline__ int row_allmax(int x) {
    x = max(x, dpp<kRowRor + 1, true>(0, x));
    x = max(x, dpp<kRowRor + 2, true>(0, x));
    x = max(x, dpp<kRowRor + 4, true>(0, x));
    x = max(x, dpp<kRowRor + 8, true>(0, x));
    return x;
}
__device__ __forceinline__ int row_allmin(int x) {
    x = min(x, dpp<kRowRor + 1, true>(0, x));
    x = min(x, dpp<kRowRor + 2, true>(0, x));
    x = min(x, dpp<kRowRor + 4, true>(0, x));
    x = min(x, dpp<kRowRor + 8, true>(0, x));
    return x;
}

// Sum over the 16 lanes of the row, result in every lane of the row.
__device__ __forceinline__ int row_allsum(int x) {
    x += dpp<kRowRor + 1, true>(0, x);
    x += dpp<kRowRor + 2, true>(0, x);
    x += dpp<kRowRor + 4, true>(0, x);
    x += dpp<kRowRor + 8, true>(0, x);
    return x;
}

// Inclusive prefix max over the row for non-negative values (lanes shifted in from outside the
// row read 0, the identity for non-negative inputs).
__device__ __forceinline__ int row_scan_max_nonneg(int x) {
    x = max(x, dpp<kRowShr + 1, true>(0, x));
    x = max(x, dpp<kRowShr + 2, true>(0, x));
    x = max(x, dpp<kRowShr + 4, true>(0, x));
    x = max(x, dpp<kRowShr + 8, true>(0, x));
    return x;
}

// Max over the whole wavefront of a value that is uniform within each 16-lane row (one value per
// group), returned as a scalar.  Row broadcasts (lane 15 of a row into the next row, then lane 31
// into rows 2 and 3) leave the maximum of all four rows in lane 63.  Each step is one max with a
// DPP operand; lanes outside the step's row mask keep their value.  The s_nop gives the DPP read
// of a register written by the previous VALU its two wait states.  The whole wavefront must be
// active.
__device__ __forceinline__ int wave_max_of_rows(int x) {
    asm volatile(
        "s_nop 1\n\t"
        "v_max_i32_dpp %0, %0, %0 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
        "s_nop 1\n\t"
        "v_max_i32_dpp %0, %0, %0 row_bcast:31 row_mask:0xc bank_mask:0xf"
        : "+v"(x));
    return __builtin_amdgcn_readlane(x, 63);
}

// Value of lane l-1 of the row; lane 0 of the row receives `fill`.
__device__ __forceinline__ int row_shift_in(int x, int fill) { return dpp<kRowShr + 1, false>(fill, x); }

// Lane layout of the groups of a wavefront.  GroupLanes<16>: a group is one 16-lane DPP row, four
// groups per wavefront.  GroupLanes<8>: two groups share each DPP row, interleaved (the group's
// parity is lane & 1, the lane within the group (lane & 15) >> 1), so that a row shift or rotation
// by an even distance stays inside one group: each primitive is the 16-lane one with its distances
// doubled and one step fewer, and eight groups share a wavefront.
template <int GL>
struct GroupLanes;

template <>
struct GroupLanes<16> {
    static constexpr int kPerWave = 4;
    __device__ static __forceinline__ int g() { return (int)(threadIdx.x & 15); }
    __device__ static __forceinline__ int gid() { return (int)(threadIdx.x >> 4); }
    // lane 0 of the calling lane's group, as a lane of the wavefront
    __device__ static __forceinline__ int leader() { return (int)(threadIdx.x & 63) & ~15; }
    __device__ static __forceinline__ int allmax(int x) { return row_allmax(x); }
    __device__ static __forceinline__ int allmin(int x) { return row_allmin(x); }
    __device__ static __forceinline__ int scan_max_nonneg(int x) { return row_scan_max_nonneg(x); }
    __device__ static __forceinline__ int shift_in(int x, int fill) { return row_shift_in(x, fill); }
    // value of the previous lane of the group; the group's first lane receives 0
    __device__ static __forceinline__ int shift_in_zero(int x) { return dpp<kRowShr + 1, true>(0, x); }
    __device__ static __forceinline__ int wave_max(int x) { return wave_max_of_rows(x); }
};

template <>
struct GroupLanes<8> {
    static constexpr int kPerWave = 8;
    __device__ static __forceinline__ int g() { return (int)((threadIdx.x & 15) >> 1); }
    __device__ static __forceinline__ int gid() { return (int)(((threadIdx.x >> 4) << 1) | (threadIdx.x & 1)); }
    __device__ static __forceinline__ int leader() { return ((int)(threadIdx.x & 63) & ~15) | (int)(threadIdx.x & 1); }
    __device__ static __forceinline__ int allmax(int x) {
        x = max(x, dpp<kRowRor + 2, true>(0, x));
        x = max(x, dpp<kRowRor + 4, true>(0, x));
        x = max(x, dpp<kRowRor + 8, true>(0, x));
        return x;
    }
    __device__ static __forceinline__ int allmin(int x) {
        x = min(x, dpp<kRowRor + 2, true>(0, x));
        x = min(x, dpp<kRowRor + 4, true>(0, x));
        x = min(x, dpp<kRowRor + 8, true>(0, x));
        return x;
    }
    // lanes 0..2d-1 of the row (the first d lanes of each group) read past the row: 0
    __device__ static __forceinline__ int scan_max_nonneg(int x) {
        x = max(x, dpp<kRowShr + 2, true>(0, x));
        x = max(x, dpp<kRowShr + 4, true>(0, x));
        x = max(x, dpp<kRowShr + 8, true>(0, x));
        return x;
    }
    __device__ static __forceinline__ int shift_in(int x, int fill) { return dpp<kRowShr + 2, false>(fill, x); }
    __device__ static __forceinline__ int shift_in_zero(int x) { return dpp<kRowShr + 2, true>(0, x); }
    // the row's two groups first (a rotation by one pairs each lane with the other group), then
    // the rows
    __device__ static __forceinline__ int wave_max(int x) {
        return wave_max_of_rows(max(x, dpp<kRowRor + 1, true>(0, x)));
    }
};

// Job slots for the 16-lane groups of a persistent grid over a queue of `count` jobs sorted by
// descending work.  With `deal`, slots below count - (groups in the grid) are dealt round-robin
// (group q of the grid takes q, q + n_groups, ...), which needs no atomic: one contended queue
// atomic per job switch stalls the whole wavefront for its round trip, a large share of a short
// job.  The last n_groups slots, the shortest jobs, come from the queue counter, so that groups
// that finish early take them and the grid's tail stays balanced.  Without `deal` (long jobs,
// where per-wavefront speed differences outweigh the atomic) every slot comes from the counter.
// Every lane of a group calls take() together.
// GroupSlotsT<8> does the same for the 8-lane groups of GroupLanes<8>.
template <int GL = 16>
struct GroupSlotsT {
    using L = GroupLanes<GL>;
    uint32_t next, stride, static_end;
    __device__ __forceinline__ GroupSlotsT(uint32_t count, bool deal) {
        const uint32_t groups_per_block = blockDim.x / GL;
        stride = gridDim.x * groups_per_block;
        next = blockIdx.x * groups_per_block + (uint32_t)L::gid();
        static_end = deal && count > stride ? count - stride : 0u;
    }
    // The group's next slot; a slot >= count means the queue is exhausted.
    __device__ __forceinline__ uint32_t take(uint32_t *head) {
        if (next < static_end) {
            const uint32_t s = next;
            next += stride;
            return s;
        }
        uint32_t s = 0;
        if (L::g() == 0) s = static_end + atomicAdd(head, 1u);
        return (uint32_t)__shfl((int)s, L::leader(), 64);
    }
};
using GroupSlots = GroupSlotsT<16>;

// Order LDS accesses of the lanes of one wavefront: everything before is visible to every lane
// of the wave after.  A wavefront's LDS operations execute in issue order, so this only has to
// stop the compiler from moving accesses across it.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace bwg
