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

// Order LDS accesses of the lanes of one wavefront: everything before is visible to every lane
// of the wave after.  A wavefront's LDS operations execute in issue order, so this only has to
// stop the compiler from moving accesses across it.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace bwg
