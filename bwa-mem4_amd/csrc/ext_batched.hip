// Batched form of the banded extension (ksw_extend2 semantics, crates/bwa-extend/src/sw.rs:203-493).
//
// Two persistent kernels, each draining a device work queue of length-binned jobs:
//
//  * ext_group_kernel<K>: one job per 16-lane DPP row ("group"), four jobs per wavefront.  Row i
//    of the DP is evaluated in parallel over the live window [beg, end): lane g owns the K
//    consecutive window columns beg + g*K .. beg + g*K + K-1 (window-relative blocking; K odd keeps
//    the 16 lanes' LDS accesses in distinct banks).  The eh arrays of the reference live in LDS over
//    absolute columns 0..qlen, so columns outside the window keep their stale values exactly as in
//    the reference.  M and E depend only on the previous row; F is the max-plus recurrence
//    F(j+1) = max(F(j) - e_ins, max(M(j) - oe_ins, 0)), evaluated in two passes: each lane runs it
//    over its block from F_in = 0, the blocks are combined by an exclusive max-scan across the row
//    (F(l) = max_{m<l} c_m - (l-1-m)*K*e_ins), and a second pass applies each lane's incoming F.
//    Row maximum with the later-column tie rule is a max over keys (h << 13 | j).  The row
//    epilogue (to-end score, z-drop, band tightening) is evaluated redundantly by the 16 lanes.
//
//  * ext_thread_kernel<QCAP>: one job per lane for short queries (qlen <= QCAP), the scalar
//    recurrence with the eh arrays in LDS laid out [column][lane].
#include <algorithm>
#include <atomic>

#include "bwg_batched_kernels.h"
#include "dpp.h"
#include "target_stream.h"

namespace bwg {
namespace {

constexpr int kKeyShift = 13;  // key = h << 13 | j, j < 8192
constexpr int kKeyMask = (1 << kKeyShift) - 1;

struct GroupJob {
    // per-job constants (identical in the 16 lanes of the group)
    TargetView t;
    int32_t qlen, tlen, h0, w;
    uint32_t out_idx;
    // row state
    int32_t i, beg, end, best, best_i, best_j, gbest_i, gscore, max_off;
    int32_t tcur;  // target code of row i
    int32_t h1base;  // h0 - (o_del + e_del*(i+1)), kept as a running difference
};

// Column state.  The reference's eh arrays hold {H[j], E[j]} per absolute column.  Wide storage
// keeps them as two int32; narrow storage packs H into the low and E into the high 16 bits of one
// word, exact whenever every value fits in 16 bits: H and E are never negative and never exceed
// h0 + qlen*max_sc, which the planner checks is below 65536 for narrow bins.  Narrow storage halves
// the block's LDS footprint, which is what limits the group kernel's occupancy.
template <bool NARROW> struct ColState;
template <> struct ColState<true> {
    using word = uint32_t;
    static __device__ __forceinline__ word pack(int32_t h, int32_t e) { return (uint32_t)h | ((uint32_t)e << 16); }
    static __device__ __forceinline__ int32_t h_of(word w) { return (int32_t)(w & 0xffffu); }
    static __device__ __forceinline__ int32_t e_of(word w) { return (int32_t)(w >> 16); }
    static __device__ __forceinline__ bool live(word w) { return w != 0; }
};
template <> struct ColState<false> {
    using word = int2;
    static __device__ __forceinline__ word pack(int32_t h, int32_t e) { return make_int2(h, e); }
    static __device__ __forceinline__ int32_t h_of(word w) { return w.x; }
    static __device__ __forceinline__ int32_t e_of(word w) { return w.y; }
    static __device__ __forceinline__ bool live(word w) { return (w.x | w.y) != 0; }
};

// Substitution score of query code `qc` against the target code whose matrix row is `row`
// (mat[t][0..4] as int8 in the low five bytes of the 64-bit pair {hi:lo}).
__device__ __forceinline__ int32_t subst(uint2 row, uint32_t qc) {
    return (int32_t)(int8_t)(uint8_t)__builtin_amdgcn_perm(row.y, row.x, qc);
}

// The substitution table of the block, written once before the persistent loop.
__device__ __forceinline__ void fill_subst_table(int8_t *tbl, const ExtScoring &sc) {
    const int t = threadIdx.x;
    if (t < 40) tbl[t] = (t & 7) < 5 ? sc.mat[(t >> 3) * 5 + (t & 7)] : (int8_t)0;
}

// Block LDS layout (bytes): the eh pairs {H[j], E[j]} of every group, (qcap+2) columns each; then
// every group's score planes, plane c holding mat[c][query[j]] for j < qcap (five int8 planes);
// then slack.  A lane always loads all K columns of its block, including columns past the window
// (their values are not used); those loads may run past the group's own region into the next
// one, and the slack keeps the last group's loads inside the allocation.
__host__ __device__ __forceinline__ int eh_bytes(int qcap) { return 8 * (qcap + 2); }
__host__ __device__ __forceinline__ int plane_bytes(int qcap) { return 5 * qcap; }
__host__ __device__ __forceinline__ size_t block_lds_bytes(int groups, int qcap, int K) {
    return (size_t)groups * (eh_bytes(qcap) + plane_bytes(qcap)) + 16 * K * 8 + 64;
}

constexpr int kGroupsPerBlock = 16;

template <int K>
__global__ __launch_bounds__(256) void ext_group_kernel(const uint8_t *__restrict__ arena,
                                                        const uint8_t *__restrict__ ref,
                                                        const ExtJob *__restrict__ jobs,
                                                        const uint32_t *__restrict__ order,
                                                        uint32_t count, uint32_t *__restrict__ head,
                                                        ExtScoring sc, int qcap,
                                                        bwg_ext_result *__restrict__ out) {
    extern __shared__ int32_t lds[];
    const int lane = threadIdx.x & 63;
    const int g = lane & 15;             // lane within the group
    const int gid = threadIdx.x >> 4;    // group within the block
    int32_t *eh = (int32_t *)((char *)lds + gid * eh_bytes(qcap));
    int8_t *planes = (int8_t *)lds + kGroupsPerBlock * eh_bytes(qcap) + gid * plane_bytes(qcap);

    const int32_t oe_del = sc.o_del + sc.e_del, oe_ins = sc.o_ins + sc.e_ins;
    const int32_t e_del = sc.e_del, e_ins = sc.e_ins;
    const int32_t Ke = K * e_ins;

    GroupJob J;
    bool active = false, exhausted = false;
    for (;;) {
        // ---- fetch and initialise a job for idle groups ------------------------------------
        if (!active && !exhausted) {
            uint32_t slot = 0;
            if (g == 0) slot = atomicAdd(head, 1u);
            slot = __shfl(slot, lane & ~15, 64);
            if (slot >= count) {
                exhausted = true;
            } else {
                const ExtJob jb = jobs[order[slot]];
                const uint8_t *q = arena + jb.q_off;
                J.t = target_view(arena, ref, jb.t_off, jb.tlen);
                J.qlen = (int32_t)jb.qlen;
                J.tlen = (int32_t)jb.tlen;
                J.h0 = jb.h0;
                J.out_idx = order[slot];
                J.w = clamp_band(sc.w, J.qlen, sc);
                // row -1 (sw.rs:263-279): H[0] = h0, H[j] = max(h0 - oe_ins - (j-1)*e_ins, 0)
                for (int j = g; j <= J.qlen + 1; j += 16) {
                    const int32_t h = j == 0 ? J.h0 : max(J.h0 - oe_ins - (j - 1) * e_ins, 0);
                    *(int2 *)(eh + 2 * j) = make_int2(j <= J.qlen ? h : 0, 0);
                }
                for (int j = g; j < J.qlen; j += 16) {
                    const int qc = q[j];
#pragma unroll
                    for (int c = 0; c < 5; ++c) planes[c * qcap + j] = sc.mat[c * 5 + qc];
                }
                J.i = 0;
                J.beg = 0;
                J.end = J.qlen;
                J.best = J.h0;
                J.best_i = J.best_j = J.gbest_i = J.gscore = -1;
                J.max_off = 0;
                J.tcur = J.tlen > 0 ? J.t[0] : 0;
                J.h1base = (int32_t)((uint32_t)J.h0 - (uint32_t)sc.o_del - (uint32_t)e_del);
                active = J.tlen > 0;
                if (!active && g == 0) {
                    bwg_ext_result r = {J.h0, 0, 0, 0, -1, 0};
                    out[J.out_idx] = r;
                }
            }
        }
        wave_lds_sync();
        // the wavefront exits once none of its four groups has work left
        if (__ballot(active || !exhausted) == 0) break;
        if (!active) continue;

        // ---- one DP row ------------------------------------------------------------------------
        const int32_t i = J.i;
        int32_t beg = J.beg, end = J.end;
        if (beg < i - J.w) beg = i - J.w;
        if (end > i + J.w + 1) end = i + J.w + 1;
        if (end > J.qlen) end = J.qlen;
        const int32_t h1_init = beg == 0 ? max(J.h1base, 0) : 0;
        const int32_t tnext = i + 1 < J.tlen ? J.t[i + 1] : 0;  // prefetched for the next row

        const int32_t base = beg + g * K;  // first window column of this lane
        const int32_t nval = min(max(end - base, 0), K);
        const int2 *ehb = (const int2 *)eh + base;
        const int8_t *sb = planes + __umul24((uint32_t)J.tcur, (uint32_t)qcap) + base;
        // pass 1: loads of all K columns are unconditional and independent, so they issue back to
        // back; columns past the window compute values that are never used
        int2 v[K];
        int32_t s[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            v[k] = ehb[k];
            s[k] = sb[k];
        }
        int32_t mm[K], gg[K];
        int32_t grun = 0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int32_t m = v[k].x != 0 ? v[k].x + s[k] : 0;
            mm[k] = m;
            gg[k] = grun;
            grun = max(max(grun - e_ins, m - oe_ins), 0);
        }
        // F entering this lane's block: exclusive max-scan of U_l = c_l + l*K*e_ins over the row
        const int32_t U = grun + g * Ke;
        const int32_t incl = row_scan_max_nonneg(U);
        const int32_t excl = row_shift_in(incl, -Ke);
        const int32_t f_in = excl - (g - 1) * Ke;

        int32_t key = 0;
        int2 *ehw = (int2 *)eh + base;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int32_t f = max(f_in - k * e_ins, gg[k]);
            const int32_t h = max(max(mm[k], v[k].y), f);
            const int32_t e = max(max(v[k].y - e_del, mm[k] - oe_del), 0);
            if (k < nval) {
                // lane-relative column; the lane's first column is added once after the loop
                key = max(key, (h << kKeyShift) | k);
                ehw[k].y = e;       // E(i+1, j)
                ehw[k + 1].x = h;   // H(i, j): the diagonal of column j+1 next row
            }
        }
        if (g == 0) {
            eh[2 * beg] = h1_init;  // H(i, beg-1)
            eh[2 * end + 1] = 0;    // E at the sentinel column
        }
        wave_lds_sync();
        key = row_allmax(nval > 0 ? key + base : 0);  // h << 13 | absolute column
        const int32_t row_max = key >> kKeyShift;
        const int32_t mj = key & kKeyMask;
        const int2 at_end = *(const int2 *)(eh + 2 * end);  // H(i, end-1) (h1_init if empty)
        const int2 at_beg = *(const int2 *)(eh + 2 * beg);
        const int32_t h1_last = at_end.x;

        bool done = false;
        const int32_t j_exit = beg < end ? end : beg;
        if (j_exit == J.qlen && J.gscore <= h1_last) {
            J.gbest_i = i;
            J.gscore = h1_last;
        }
        if (row_max == 0) {
            done = true;
        } else if (row_max > J.best) {
            J.best = row_max;
            J.best_i = i;
            J.best_j = mj;
            J.max_off = max(J.max_off, abs(mj - i));
        } else if (sc.zdrop > 0) {
            const int32_t di = i - J.best_i, dj = mj - J.best_j;
            const int32_t dd = di > dj ? di - dj : dj - di, ge = di > dj ? e_del : e_ins;
            const int32_t gap = sc.small_ext ? __mul24(dd, ge) : dd * ge;
            if (J.best - row_max - gap > sc.zdrop) done = true;
        }
        if (!done) {
            // band tightening (sw.rs:464-480): first live column in [beg, end), last live column
            // scanning down from end.  The edges are usually live, so test them first.
            int32_t nb = beg;
            if (beg < end && (at_beg.x | at_beg.y) == 0) {
                for (;;) {
                    const int j = nb + g;
                    bool live = false;
                    if (j < end) {
                        const int2 x = *(const int2 *)(eh + 2 * j);
                        live = (x.x | x.y) != 0;
                    }
                    const int32_t cand = row_allmin(live ? j : 0x7fffffff);
                    if (cand != 0x7fffffff) { nb = cand; break; }
                    nb += 16;
                    if (nb >= end) { nb = end; break; }
                }
            }
            int32_t ne = end;
            if ((at_end.x | at_end.y) == 0) {
                for (;;) {
                    const int j = ne - g;
                    bool live = false;
                    if (j >= nb) {
                        const int2 x = *(const int2 *)(eh + 2 * j);
                        live = (x.x | x.y) != 0;
                    }
                    const int32_t cand = row_allmax(live ? j : -1);
                    if (cand >= 0) { ne = cand; break; }
                    ne -= 16;
                    if (ne < nb) { ne = nb - 1; break; }
                }
            }
            J.beg = nb;
            J.end = ne + 2 < J.qlen ? ne + 2 : J.qlen;
            J.i = i + 1;
            J.tcur = tnext;
            J.h1base = (int32_t)((uint32_t)J.h1base - (uint32_t)e_del);
            if (J.i >= J.tlen) done = true;
        }
        if (done) {
            if (g == 0) {
                bwg_ext_result r;
                r.score = J.best;
                r.qle = J.best_j + 1;
                r.tle = J.best_i + 1;
                r.gtle = J.gbest_i + 1;
                r.gscore = J.gscore;
                r.max_off = J.max_off;
                out[J.out_idx] = r;
            }
            active = false;
        }
    }
}

// ---- short queries: one job per lane --------------------------------------------------------
// Column state and query codes in LDS laid out [column][lane], so the 64 lanes of a wavefront
// touch consecutive words; narrow storage as in the group kernel.
template <int QCAP, bool NARROW>
__global__ __launch_bounds__(64) void ext_thread_kernel(const uint8_t *__restrict__ arena,
                                                        const uint8_t *__restrict__ ref,
                                                        const ExtJob *__restrict__ jobs,
                                                        const uint32_t *__restrict__ order,
                                                        uint32_t count, uint32_t *__restrict__ head,
                                                        ExtScoring sc,
                                                        bwg_ext_result *__restrict__ out) {
    using CS = ColState<NARROW>;
    using W = typename CS::word;
    __shared__ W eh[(QCAP + 1) * 64];
    __shared__ uint8_t codes[QCAP * 64];
    __shared__ __align__(8) int8_t tbl[64];
    fill_subst_table(tbl, sc);
    __syncthreads();
    const int lane = threadIdx.x;
    const int32_t oe_del = sc.o_del + sc.e_del, oe_ins = sc.o_ins + sc.e_ins;
    const int32_t e_del = sc.e_del, e_ins = sc.e_ins;
    for (;;) {
        uint32_t slot = 0;
        if (lane == 0) slot = atomicAdd(head, 64u);
        slot = __shfl(slot, 0, 64);
        if (slot >= count) break;
        const uint32_t my = slot + lane;
        if (my < count) {
            const uint32_t idx = order[my];
            const ExtJob jb = jobs[idx];
            const uint8_t *q = arena + jb.q_off;
            const TargetView t = target_view(arena, ref, jb.t_off, jb.tlen);
            const int32_t qlen = (int32_t)jb.qlen, tlen = (int32_t)jb.tlen, h0 = jb.h0;
            const int32_t w = clamp_band(sc.w, qlen, sc);
            for (int j = 0; j <= qlen; ++j)
                eh[j * 64 + lane] = CS::pack(j == 0 ? h0 : max(h0 - oe_ins - (j - 1) * e_ins, 0), 0);
            for (int j = 0; j < qlen; ++j) codes[j * 64 + lane] = q[j];
            int32_t best = h0, best_i = -1, best_j = -1, gbest_i = -1, gscore = -1, max_off = 0;
            int32_t beg = 0, end = qlen;
            LaneTargetStream ts;
            if (tlen > 0) ts.start(t, tlen);
            for (int32_t i = 0; i < tlen; ++i) {
                const uint2 trow = *(const uint2 *)(tbl + 8 * ts.at(i, t, tlen));
                int32_t f = 0, key = 0;
                if (beg < i - w) beg = i - w;
                if (end > i + w + 1) end = i + w + 1;
                if (end > qlen) end = qlen;
                int32_t h1 = beg == 0 ? max(h0 - (sc.o_del + e_del * (i + 1)), 0) : 0;
                int32_t j = beg;
                for (; j < end; ++j) {
                    const W v = eh[j * 64 + lane];
                    const int32_t hd = CS::h_of(v), ev = CS::e_of(v);
                    const int32_t m = hd != 0 ? hd + subst(trow, codes[j * 64 + lane]) : 0;
                    const int32_t h = max(max(m, ev), f);
                    const int32_t e = max(max(ev - e_del, m - oe_del), 0);
                    eh[j * 64 + lane] = CS::pack(h1, e);
                    h1 = h;
                    key = max(key, (h << kKeyShift) | j);
                    f = max(f - e_ins, max(m - oe_ins, 0));
                }
                eh[end * 64 + lane] = CS::pack(h1, 0);
                const int32_t row_max = key >> kKeyShift, mj = key & kKeyMask;
                if (j == qlen && gscore <= h1) {
                    gbest_i = i;
                    gscore = h1;
                }
                if (row_max == 0) break;
                if (row_max > best) {
                    best = row_max;
                    best_i = i;
                    best_j = mj;
                    max_off = max(max_off, abs(mj - i));
                } else if (sc.zdrop > 0) {
                    const int32_t di = i - best_i, dj = mj - best_j;
                    if (di > dj) {
                        if (best - row_max - (di - dj) * e_del > sc.zdrop) break;
                    } else if (best - row_max - (dj - di) * e_ins > sc.zdrop) {
                        break;
                    }
                }
                int32_t a = beg;
                while (a < end && !CS::live(eh[a * 64 + lane])) ++a;
                beg = a;
                int32_t z = end;
                while (z >= beg && !CS::live(eh[z * 64 + lane])) --z;
                end = z + 2 < qlen ? z + 2 : qlen;
            }
            bwg_ext_result r;
            r.score = best;
            r.qle = best_j + 1;
            r.tle = best_i + 1;
            r.gtle = gbest_i + 1;
            r.gscore = gscore;
            r.max_off = max_off;
            out[idx] = r;
        }
    }
}

template <int K>
hipError_t launch_group(hipStream_t st, const uint8_t *arena, const uint8_t *ref, const ExtJob *jobs, const uint32_t *order,
                        uint32_t count, uint32_t *head, const ExtScoring &sc, int qcap,
                        bwg_ext_result *out, int grid) {
    const size_t lds = block_lds_bytes(kGroupsPerBlock, qcap, K);
    hipLaunchKernelGGL(ext_group_kernel<K>, dim3(grid), dim3(256), lds, st, arena, ref, jobs, order, count,
                       head, sc, qcap, out);
    return hipGetLastError();
}

// Occupancy queries are cached: they depend only on the kernel and its LDS size.
template <int K>
int group_occupancy(int qcap) {
    static std::atomic<int> cache[130];
    const int slot = std::min(qcap / 32, 129);
    int v = cache[slot].load(std::memory_order_relaxed);
    if (v > 0) return v;
    int blocks = 0;
    const size_t lds = block_lds_bytes(kGroupsPerBlock, qcap, K);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, ext_group_kernel<K>, 256, lds) != hipSuccess)
        blocks = 1;
    v = blocks > 0 ? blocks : 1;
    cache[slot].store(v, std::memory_order_relaxed);
    return v;
}

template <int QCAP, bool NARROW>
int thread_occupancy() {
    static std::atomic<int> cache{0};
    int v = cache.load(std::memory_order_relaxed);
    if (v > 0) return v;
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, ext_thread_kernel<QCAP, NARROW>, 64, 0) != hipSuccess)
        blocks = 1;
    v = blocks > 0 ? blocks : 1;
    cache.store(v, std::memory_order_relaxed);
    return v;
}

}  // namespace

size_t ext_group_lds_bytes(int K, int qcap) { return block_lds_bytes(kGroupsPerBlock, qcap, K); }

int ext_group_blocks_per_cu(int K, int qcap) {
    switch (K) {
#define BWG_K(k) case k: return group_occupancy<k>(qcap);
        BWG_EXT_GROUP_KS
#undef BWG_K
    default: return 1;
    }
}

hipError_t launch_ext_group(hipStream_t st, int K, const uint8_t *arena, const uint8_t *ref, const ExtJob *jobs,
                            const uint32_t *order, uint32_t count, uint32_t *head,
                            const ExtScoring &sc, int qcap, bwg_ext_result *out, int grid) {
    switch (K) {
#define BWG_K(k) case k: return launch_group<k>(st, arena, ref, jobs, order, count, head, sc, qcap, out, grid);
        BWG_EXT_GROUP_KS
#undef BWG_K
    default: return hipErrorInvalidValue;
    }
}

int ext_thread_blocks_per_cu(int qcap, bool narrow) {
    switch (qcap) {
    case 16: return narrow ? thread_occupancy<16, true>() : thread_occupancy<16, false>();
    case 32: return narrow ? thread_occupancy<32, true>() : thread_occupancy<32, false>();
    default: return 1;
    }
}

hipError_t launch_ext_thread(hipStream_t st, int qcap, bool narrow, const uint8_t *arena, const uint8_t *ref,
                             const ExtJob *jobs, const uint32_t *order, uint32_t count, uint32_t *head,
                             const ExtScoring &sc, bwg_ext_result *out, int grid) {
#define BWG_T(Q, N)                                                                                        \
    hipLaunchKernelGGL((ext_thread_kernel<Q, N>), dim3(grid), dim3(64), 0, st, arena, ref, jobs, order, count, \
                       head, sc, out)
    if (qcap == 16 && narrow) BWG_T(16, true);
    else if (qcap == 16) BWG_T(16, false);
    else if (qcap == 32 && narrow) BWG_T(32, true);
    else if (qcap == 32) BWG_T(32, false);
    else return hipErrorInvalidValue;
#undef BWG_T
    return hipGetLastError();
}

}  // namespace bwg
