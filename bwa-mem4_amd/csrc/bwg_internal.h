// Internal declarations shared by the host runtime (bwg_api.cpp) and the kernel files.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "../../include/bwg.h"

namespace bwg {

// Target location: t_off indexes the batch arena, or, with kTFromRef set, a device-resident
// reference; with kTReverse set the target is that window read back to front (a left extension's
// target, or a reverse-strand window), i.e. target[i] = base[off + tlen - 1 - i].
constexpr uint64_t kTFromRef = 1ull << 62;
constexpr uint64_t kTReverse = 1ull << 63;
constexpr uint64_t kTOffMask = kTFromRef - 1;

struct TargetView {
    const uint8_t *p;
    int32_t step;
    __host__ __device__ uint8_t operator[](int32_t i) const { return p[(ptrdiff_t)i * step]; }
};
__host__ __device__ inline TargetView target_view(const uint8_t *arena, const uint8_t *ref,
                                                  uint64_t t_off, uint32_t tlen) {
    const uint8_t *base = (t_off & kTFromRef) ? ref : arena;
    const uint64_t o = t_off & kTOffMask;
    TargetView v;
    if (t_off & kTReverse) {
        v.p = base + o + (tlen ? tlen - 1 : 0);
        v.step = -1;
    } else {
        v.p = base + o;
        v.step = 1;
    }
    return v;
}

// Job descriptor as the kernels read it: 32 bytes, one load per job.
struct ExtJob {
    uint64_t q_off, t_off;  // q_off into the device arena; t_off as described above
    uint32_t qlen, tlen;
    int32_t h0;
    uint32_t aux;  // plain form: scratch slot offset in int2 units / 8; batched: original index
};
struct KswJob {
    uint64_t q_off, t_off;
    uint32_t qlen, tlen;
    uint32_t aux0, aux1;  // batched: row-maximum tile and lane; plain: scratch offset / 16 in aux0
};
// First row-maximum slot of a batched rescue job: tiles of 64 jobs, row-major within a tile, so a
// job's slot for row i is row_slot + 64*i.  aux0 counts 64-word row blocks, aux1 is the lane.
__host__ __device__ inline size_t row_slot(const KswJob &j) { return ((size_t)j.aux0 << 6) | j.aux1; }

// Scoring shared by all jobs of an extension batch, plus the values derived from it once.
struct ExtScoring {
    int8_t mat[28];
    int32_t o_del, e_del, o_ins, e_ins, w, end_bonus, zdrop;
    int32_t max_sc;
    int32_t small_ext;  // e_del, e_ins < 2^14: gap products of row/column distances fit 24-bit multiplies
};
struct KswScoring {
    int8_t mat[28];
    int32_t o_del, e_del, o_ins, e_ins, minsc, max_sc;
};

inline int32_t mat_max(const int8_t *mat) {
    int32_t mx = mat[0];
    for (int k = 1; k < 25; ++k) mx = mat[k] > mx ? mat[k] : mx;
    return mx;
}

inline ExtScoring ext_scoring(const bwg_ext_params &p) {
    ExtScoring sc{};
    for (int k = 0; k < 25; ++k) sc.mat[k] = p.mat[k];
    sc.o_del = p.o_del;
    sc.e_del = p.e_del;
    sc.o_ins = p.o_ins;
    sc.e_ins = p.e_ins;
    sc.w = p.w;
    sc.end_bonus = p.end_bonus;
    sc.zdrop = p.zdrop;
    sc.max_sc = mat_max(p.mat);
    sc.small_ext = p.e_del < (1 << 14) && p.e_ins < (1 << 14);
    return sc;
}

// The ungapped fast path's preconditions (crates/bwa-neon/src/batched.rs:493-521, 756-764): a
// matrix of the uniform DNA shape, and the largest mismatch count for which the ungapped diagonal
// strictly beats the cheapest gap, X * (a + b) < o_min + e_min (-1: never).
inline bool is_uniform_dna(const ExtScoring &sc) {
    const int8_t a = sc.mat[0], mm = sc.mat[1], npen = sc.mat[4];
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j) {
            const int8_t want = (i == 4 || j == 4) ? npen : (i == j ? a : mm);
            if (sc.mat[i * 5 + j] != want) return false;
        }
    return true;
}
inline int32_t ungapped_xmax(const ExtScoring &sc) {
    const int32_t a = sc.mat[0], b = -(int32_t)sc.mat[1];
    const int32_t o_min = sc.o_del < sc.o_ins ? sc.o_del : sc.o_ins;
    const int32_t e_min = sc.e_del < sc.e_ins ? sc.e_del : sc.e_ins;
    return a + b > 0 ? (o_min + e_min - 1) / (a + b) : -1;
}

inline KswScoring ksw_scoring(const bwg_ksw_params &p) {
    KswScoring sc{};
    for (int k = 0; k < 25; ++k) sc.mat[k] = p.mat[k];
    sc.o_del = p.o_del;
    sc.e_del = p.e_del;
    sc.o_ins = p.o_ins;
    sc.e_ins = p.e_ins;
    sc.minsc = p.minsc;
    sc.max_sc = p.max_sc;
    return sc;
}

// Integer closed form of ksw_extend2's f64 band clamp (crates/bwa-extend/src/sw.rs:288-308):
// trunc((double)n/e + 1.0) floored at 1 equals (n <= 0 ? 1 : n/e + 1) for n = qlen*max_sc +
// end_bonus - o and e >= 1.
__host__ __device__ inline int32_t clamp_band(int32_t w, int32_t qlen, const ExtScoring &s) {
    int64_t base = (int64_t)qlen * s.max_sc + s.end_bonus;
    int64_t ni = base - s.o_ins, nd = base - s.o_del;
    // With qlen < 2^24, max_sc < 128 and bwa's non-negative penalties a positive numerator fits 32
    // unsigned bits, and the quotient is a 32-bit division (a 64-bit one costs the GPU about a
    // hundred instructions at every job start); the result is the 64-bit expression's either way
    constexpr int64_t k32 = (int64_t)1 << 32;
    int32_t a = ni <= 0 ? 1
                : ni < k32 ? (int32_t)((int64_t)((uint32_t)ni / (uint32_t)s.e_ins) + 1)
                           : (int32_t)(ni / s.e_ins + 1);
    int32_t b = nd <= 0 ? 1
                : nd < k32 ? (int32_t)((int64_t)((uint32_t)nd / (uint32_t)s.e_del) + 1)
                           : (int32_t)(nd / s.e_del + 1);
    if (w > a) w = a;
    if (w > b) w = b;
    return w;
}

// bwa's rescue lane rule (crates/bwa-mem/src/pe.rs:420-424, matesw.rs:676-686).
__host__ __device__ inline int32_t ksw_lanes(int32_t qlen, int32_t max_sc) {
    return qlen * max_sc < 250 ? 16 : 8;
}
__host__ __device__ inline int32_t ksw_padded(int32_t qlen, int32_t max_sc) {
    int32_t l = ksw_lanes(qlen, max_sc);
    return (qlen + l - 1) / l * l;
}

// ---- launchers (defined in the .hip files) ----
// Plain form: one thread per job, state in global memory.
hipError_t launch_ext_plain(hipStream_t st, const uint8_t *arena, const uint8_t *ref, const ExtJob *jobs,
                            uint32_t n, const ExtScoring &sc, int2 *scratch, bwg_ext_result *out);
hipError_t launch_ksw_plain(hipStream_t st, const uint8_t *arena, const uint8_t *ref, const KswJob *jobs,
                            uint32_t n, const KswScoring &sc, uint8_t *scratch, bwg_ksw_result *out);

// Same kernels over a compacted job list whose k-th result goes to out[out_index[k]]; with
// `active` set, only jobs whose flag active[out_index[k]] is non-zero run.
hipError_t launch_ext_plain_scatter(hipStream_t st, const uint8_t *arena, const uint8_t *ref, const ExtJob *jobs,
                                    const uint32_t *out_index, uint32_t n, const ExtScoring &sc,
                                    int2 *scratch, bwg_ext_result *out, const uint8_t *active = nullptr);
hipError_t launch_ksw_plain_scatter(hipStream_t st, const uint8_t *arena, const uint8_t *ref, const KswJob *jobs,
                                    const uint32_t *out_index, uint32_t n, const KswScoring &sc,
                                    uint8_t *scratch, bwg_ksw_result *out);

// Device-to-device copy of `bytes` (a multiple of 16) over `blocks` workgroups of 256 threads.
hipError_t launch_hbm_copy(hipStream_t st, const void *src, void *dst, size_t bytes, uint32_t blocks);

}  // namespace bwg
