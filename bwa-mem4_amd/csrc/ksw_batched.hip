// Batched form of mate rescue (batched_ksw_align2 semantics, crates/bwa-neon/src/matesw.rs:299-497;
// per-pass recurrence of ksw_local_fwd, crates/bwa-extend/src/sw.rs:921-1075).
//
// ksw_group_kernel<K, REV>: one job per 16-lane DPP row, four jobs per wavefront, persistent over a
// device work queue.  Lane g owns the K padded query columns g*K .. g*K+K-1 and keeps H(i-1, .) and
// E(i, .) of those columns in registers; there is no band, so the whole padded row is live every
// row.  Within a row the gap-in-target recurrence F(j+1) = max(F(j)-e_ins, H(j)-oe_ins, 0) with
// H(j) = max(D(j), F(j)), D(j) = max(H(i-1,j-1)+s, E(i,j), 0) reduces, because o_ins >= 0 makes
// F(j)-oe_ins <= F(j)-e_ins, to F(j+1) = max(F(j)-e_ins, D(j)-oe_ins, 0): a max-plus recurrence
// over D, which depends on the previous row only.  It is evaluated as in the extension kernel: a
// per-lane pass from F = 0, an exclusive max-scan across the row, and a second pass that applies
// the incoming F.  The row maximum with the first-column tie rule is a max over keys
// (h << 13 | 8191-j).
//
// The forward pass (REV = false) records each row's maximum for the suboptimal tracker and the
// first row/column reaching the best score; ksw_tracker_kernel then replays the row maxima through
// the tracker.  The reverse start-recovery pass (REV = true) runs on the reversed prefixes
// query[..=qe], target[..=te], read back to front in place, padded by its own lane rule, and stops
// at the first row whose maximum reaches the forward score.
#include <atomic>

#include "bwg_batched_kernels.h"
#include "dpp.h"

namespace bwg {
namespace {

constexpr int kKeyShift = 13;
constexpr int kKeyMask = (1 << kKeyShift) - 1;

template <int K, bool REV>
__global__ __launch_bounds__(256) void ksw_group_kernel(const uint8_t *__restrict__ arena,
                                                        const uint8_t *__restrict__ ref,
                                                        const KswJob *__restrict__ jobs,
                                                        const uint32_t *__restrict__ order,
                                                        uint32_t count, uint32_t *__restrict__ head,
                                                        KswScoring sc, bwg_ksw_result *__restrict__ out,
                                                        int32_t *__restrict__ rowmax) {
    // column q of the matrix for target codes 0..4 at tblT[8*q]: one 4-byte read gives the
    // profile word of a query base, the fifth byte its score against N
    __shared__ __align__(8) uint8_t tblT[40];
    if (threadIdx.x < 40) {
        const int qc = threadIdx.x >> 3, tc = threadIdx.x & 7;
        tblT[threadIdx.x] = tc < 5 ? (uint8_t)sc.mat[tc * 5 + qc] : (uint8_t)0;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int g = lane & 15;
    const int32_t oe_del = sc.o_del + sc.e_del, oe_ins = sc.o_ins + sc.e_ins;
    const int32_t e_del = sc.e_del, e_ins = sc.e_ins;
    const int32_t Ke = K * e_ins;

    bool active = false, exhausted = false;
    // per-job state
    TargetView t = {nullptr, 1};
    int32_t tlen = 0, endsc = 0, i = 0, gmax = 0, te = -1, qe = -1, tnext = 0;
    uint32_t idx = 0;
    int32_t *rm = nullptr;
    uint32_t prof[K], p4w[(K + 3) / 4];
    int32_t km[K], kc[K], hrow[K], erow[K];

    GroupSlots slots(count, REV);  // the forward pass's long jobs balance better dynamically
    for (;;) {
        if (!active && !exhausted) {
            const uint32_t slot = slots.take(head);
            if (slot >= count) {
                exhausted = true;
            } else {
                idx = order[slot];
                const KswJob jb = jobs[idx];
                const uint8_t *q = arena + jb.q_off;
                t = target_view(arena, ref, jb.t_off, jb.tlen);
                int32_t qlen = (int32_t)jb.qlen;
                tlen = (int32_t)jb.tlen;
                endsc = 0x7fffffff;
                bool run = qlen > 0 && tlen > 0;
                if (REV) {
                    const bwg_ksw_result f = out[idx];
                    run = f.score >= sc.minsc && f.qe >= 0;
                    qlen = min(f.qe + 1, qlen);
                    tlen = f.te + 1;
                    endsc = f.score;
                    te = f.te;  // forward coordinates, kept for the start computation
                    qe = f.qe;
                } else {
                    rm = rowmax + row_slot(jb);
                    if (!run && g == 0) {
                        bwg_ksw_result r;
                        r.score = 0;
                        r.qb = r.qe = r.tb = r.te = r.score2 = r.te2 = -1;
                        out[idx] = r;
                    }
                }
                if (run) {
                    const int32_t qpad = ksw_padded(qlen, sc.max_sc);
#pragma unroll
                    for (int k = 0; k < K; ++k) {
                        const int j = g * K + k;
                        uint32_t p = 0;
                        int32_t n4 = 0;
                        if (j < qlen) {
                            const int qc = REV ? q[qlen - 1 - j] : q[j];
                            p = *(const uint32_t *)(tblT + 8 * qc);
                            n4 = tblT[8 * qc + 4];
                        }
                        prof[k] = p;  // padding columns score 0 against every target base
                        if ((k & 3) == 0) p4w[k >> 2] = 0;
                        p4w[k >> 2] |= (uint32_t)n4 << (8 * (k & 3));
                        km[k] = j < qpad ? (1 << kKeyShift) : 0;  // columns past the padded query are
                        kc[k] = j < qpad ? kKeyMask - j : 0;      // not part of the DP: keyed out
                        hrow[k] = 0;
                        erow[k] = 0;
                    }
                    i = 0;
                    gmax = 0;
                    if (!REV) {
                        te = -1;
                        qe = -1;
                    }
                    tnext = REV ? t[tlen - 1] : t[0];
                    active = true;
                }
            }
        }
        if (__ballot(active || !exhausted) == 0) break;
        if (!active) continue;

        // ---- one DP row ----------------------------------------------------------------------
        const int32_t tc = tnext;
        if (i + 1 < tlen) tnext = REV ? t[tlen - 2 - i] : t[i + 1];
        const bool trow_n = tc >= 4;
        const uint32_t tsh = (uint32_t)tc * 8u;
        int32_t hd = row_shift_in(hrow[K - 1], 0);  // H(i-1, g*K-1); 0 left of column 0
        int32_t dd[K], gg[K];
        int32_t grun = 0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int32_t s = trow_n ? (int32_t)(int8_t)(p4w[k >> 2] >> (8 * (k & 3)))
                                     : (int32_t)__builtin_amdgcn_sbfe(prof[k], tsh, 8);
            const int32_t d = max(max(hd + s, erow[k]), 0);
            hd = hrow[k];
            dd[k] = d;
            gg[k] = grun;
            grun = max(max(grun - e_ins, d - oe_ins), 0);
        }
        const int32_t U = grun + g * Ke;
        const int32_t incl = row_scan_max_nonneg(U);
        const int32_t excl = row_shift_in(incl, -Ke);
        const int32_t f_in = excl - (g - 1) * Ke;
        int32_t key = 0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int32_t f = max(f_in - k * e_ins, gg[k]);
            const int32_t h = max(dd[k], f);
            key = max(key, (int32_t)(__umul24((uint32_t)h, (uint32_t)km[k]) + (uint32_t)kc[k]));
            erow[k] = max(max(erow[k] - e_del, h - oe_del), 0);
            hrow[k] = h;
        }
        key = row_allmax(key);
        const int32_t imax = key >> kKeyShift;
        if (!REV && __builtin_expect(g == 0, 1)) *rm = imax;  // under the exec mask, no branch
        if (!REV) rm += 64;  // row i+1's slot (a running pointer: one 64-bit add a row)
        bool done = false;
        if (imax > gmax) {
            gmax = imax;
            if (REV) {
                if (gmax >= endsc) {
                    // first row reaching the forward score: the alignment starts rte/rqe back
                    if (g == 0 && gmax == endsc) {
                        const int32_t rqe = kKeyMask - (key & kKeyMask);
                        out[idx].tb = te - i;
                        out[idx].qb = qe - rqe;
                    }
                    done = true;
                }
            } else {
                te = i;
                qe = kKeyMask - (key & kKeyMask);
            }
        }
        ++i;
        if (i >= tlen) done = true;
        if (done) {
            if (!REV && g == 0) {
                bwg_ksw_result r;
                r.score = gmax;
                r.qb = -1;
                r.qe = te >= 0 ? qe : -1;
                r.tb = -1;
                r.te = te;
                r.score2 = -1;
                r.te2 = -1;
                out[idx] = r;
            }
            active = false;
        }
    }
}

// ---- packed 16-bit rescue pass ----------------------------------------------------------------
// The same DP as ksw_group_kernel with two query columns per 32-bit lane register: lane g owns the
// K = 2*KH padded columns g*K .. g*K+K-1 as KH pairs (g*K + p, g*K + KH + p), so every VALU op of
// the cell update (v_pk_add/sub/max_i16) advances both halves of the lane's block.  Exact when every
// score fits 16 bits, which the planner guarantees by routing here only jobs with qlen times the
// largest matrix score below 500 and gap penalties (open + extend) below 2^14: H stays below 500,
// so H*64 + 63 fits, and E and F are kept without their floor at 0 (every reader applies that floor
// through a max with a non-negative value), so they stay above -2*(o + e).
//
// The row maximum with the first-column tie rule is a 16-bit key h*64 + (63 - c) per half, c the
// column within the lane's block, -32768 added for columns past the padded query so they never win;
// a lane's best key k16 becomes the 32-bit key h << 10 | (15-g) << 6 | (63-c), ordered by the row
// maximum, then the lowest lane, then the lowest column in the lane: the lowest query column, as
// lanes own consecutive column blocks.  It is k16 + (k16 >> 6)*960 + (15-g)*64, one shift, one
// multiply-add and one add.  F enters the block's upper half
// as max(F_in - KH*e_ins, c_lo), c_lo the lower half's exit from F = 0.  The query profile (five
// target codes x KH column pairs) sits in LDS, lane-major, one 4-byte read per pair and row.
typedef short pk16 __attribute__((ext_vector_type(2)));

// Query column of a row maximum's 32-bit key h << 10 | (GL-1-g) << 6 | (63-c): lane g of a GL-lane
// group owns columns g*K .. g*K+K-1.
template <int GL>
__device__ __forceinline__ int32_t pk_key_column(int32_t key, int K) {
    return (GL - 1 - ((key >> 6) & 15)) * K + 63 - (key & 63);
}

__device__ __forceinline__ pk16 pk_splat(int32_t v) { return (pk16){(short)v, (short)v}; }
__device__ __forceinline__ pk16 pk_max(pk16 a, pk16 b) { return __builtin_elementwise_max(a, b); }
__device__ __forceinline__ uint32_t pk_bits(pk16 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ pk16 pk_from(uint32_t w) { return __builtin_bit_cast(pk16, w); }
// h*64 + k in each half, one v_pk_mad_u16 (the compiler otherwise emits a shift and an add); the
// low 16 bits of the product and sum are the same for signed and unsigned halves.  op_sel_hi 0 on
// the constant makes the upper half read the constant's low 16 bits (64) as well.
__device__ __forceinline__ pk16 pk_key(pk16 h, pk16 k) {
    uint32_t r;
    asm("v_pk_mad_u16 %0, %1, 64, %2 op_sel_hi:[1,0,1]" : "=v"(r) : "v"(pk_bits(h)), "v"(pk_bits(k)));
    return pk_from(r);
}

// GL = 8: the same pass with 8-lane groups (GroupLanes<8>, two per DPP row), eight jobs a
// wavefront and twice the columns per lane for the same query, so that the row's fixed work (the
// profile reads, the scan, the row maximum and its bookkeeping) is shared by twice the cells.
template <int KH, bool REV, int GL>
__global__ __launch_bounds__(256) void ksw_pk_kernel(const uint8_t *__restrict__ arena,
                                                     const uint8_t *__restrict__ ref,
                                                     const KswJob *__restrict__ jobs,
                                                     const uint32_t *__restrict__ order,
                                                     uint32_t count, uint32_t *__restrict__ head,
                                                     KswScoring sc, bwg_ksw_result *__restrict__ out,
                                                     int32_t *__restrict__ rowmax) {
    constexpr int K = 2 * KH;
    __shared__ __align__(8) uint8_t tblT[40];
    __shared__ uint32_t prof[5 * KH * 256];  // [target code][pair][thread]
    if (threadIdx.x < 40) {
        const int qc = threadIdx.x >> 3, tc = threadIdx.x & 7;
        tblT[threadIdx.x] = tc < 5 ? (uint8_t)sc.mat[tc * 5 + qc] : (uint8_t)0;
    }
    __syncthreads();
    using L = GroupLanes<GL>;
    const int g = L::g();
    const int32_t oe_del = sc.o_del + sc.e_del, oe_ins = sc.o_ins + sc.e_ins;
    const int32_t e_del = sc.e_del, e_ins = sc.e_ins;
    const int32_t Ke = K * e_ins, KHe = KH * e_ins;
    const uint32_t g64 = (uint32_t)(GL - 1 - g) * 64u;  // this lane's field of the row-maximum key
    const pk16 e_del2 = pk_splat(e_del), oe_del2 = pk_splat(oe_del), zero2 = pk_splat(0);

    bool active = false, exhausted = false;
    TargetView t = {nullptr, 1};
    int32_t tlen = 0, endsc = 0, i = 0, gmax = 0, te = -1, qe = -1, tnext = 0;
    uint32_t idx = 0;
    int32_t *rm = nullptr;
    pk16 hrow[KH], erow[KH], kk[KH];

    GroupSlotsT<GL> slots(count, REV);  // the forward pass's long jobs balance better dynamically
    for (;;) {
        if (!active && !exhausted) {
            const uint32_t slot = slots.take(head);
            if (slot >= count) {
                exhausted = true;
            } else {
                idx = order[slot];
                const KswJob jb = jobs[idx];
                const uint8_t *q = arena + jb.q_off;
                t = target_view(arena, ref, jb.t_off, jb.tlen);
                int32_t qlen = (int32_t)jb.qlen;
                tlen = (int32_t)jb.tlen;
                endsc = 0x7fffffff;
                bool run = qlen > 0 && tlen > 0;
                if (REV) {
                    const bwg_ksw_result f = out[idx];
                    run = f.score >= sc.minsc && f.qe >= 0;
                    qlen = min(f.qe + 1, qlen);
                    tlen = f.te + 1;
                    endsc = f.score;
                    te = f.te;  // forward coordinates, kept for the start computation
                    qe = f.qe;
                } else {
                    rm = rowmax + row_slot(jb);
                    if (!run && g == 0) {
                        bwg_ksw_result r;
                        r.score = 0;
                        r.qb = r.qe = r.tb = r.te = r.score2 = r.te2 = -1;
                        out[idx] = r;
                    }
                }
                if (run) {
                    const int32_t qpad = ksw_padded(qlen, sc.max_sc);
#pragma unroll
                    for (int p = 0; p < KH; ++p) {
                        const int jl = g * K + p, jh = g * K + KH + p;
                        // padding columns score 0 against every target base
                        const uint32_t ql = jl < qlen ? (REV ? q[qlen - 1 - jl] : q[jl]) : 8u;
                        const uint32_t qh = jh < qlen ? (REV ? q[qlen - 1 - jh] : q[jh]) : 8u;
#pragma unroll
                        for (int c = 0; c < 5; ++c) {
                            const int32_t sl = ql < 8u ? (int32_t)(int8_t)tblT[8 * ql + c] : 0;
                            const int32_t sh = qh < 8u ? (int32_t)(int8_t)tblT[8 * qh + c] : 0;
                            prof[(c * KH + p) * 256 + threadIdx.x] = pk_bits((pk16){(short)sl, (short)sh});
                        }
                        // columns past the padded query are not part of the DP: keyed out
                        kk[p] = (pk16){(short)(jl < qpad ? 63 - p : -32768), (short)(jh < qpad ? 63 - KH - p : -32768)};
                        hrow[p] = zero2;
                        erow[p] = zero2;
                    }
                    i = 0;
                    gmax = 0;
                    if (!REV) {
                        te = -1;
                        qe = -1;
                    }
                    tnext = REV ? t[tlen - 1] : t[0];
                    active = true;
                }
            }
        }
        if (__ballot(active || !exhausted) == 0) break;
        if (!active) continue;

        // ---- one DP row ----------------------------------------------------------------------
        const int32_t tc = tnext;
        if (i + 1 < tlen) tnext = REV ? t[tlen - 2 - i] : t[i + 1];
        const uint32_t *pr = prof + __umul24((uint32_t)tc, (uint32_t)(KH * 256)) + threadIdx.x;
        pk16 s2[KH];
#pragma unroll
        for (int p = 0; p < KH; ++p) s2[p] = pk_from(pr[p * 256]);
        // H(i-1, column-1): lower half of pair 0 from the previous lane's last column (0 left of
        // column 0), upper half of pair 0 from this lane's last lower column, then pair p-1
        // ({this lane's last pair, the previous lane's last pair} >> 16): one funnel shift
        const uint32_t prev_pair = (uint32_t)L::shift_in((int32_t)pk_bits(hrow[KH - 1]), 0);
        pk16 hd = pk_from(__builtin_amdgcn_alignbit(pk_bits(hrow[KH - 1]), prev_pair, 16u));
        // F along each half carried with its drift G_p = F_p + p*e_ins: F_{p+1} = max(F_p - e_ins,
        // M_p - oe_ins) becomes G_{p+1} = max(G_p, M_p + (p+1)*e_ins - oe_ins), one op per pair
        pk16 dd[KH], dm[KH];
        pk16 grun = zero2;
#pragma unroll
        for (int p = 0; p < KH; ++p) {
            const pk16 d = pk_max(pk_max(hd + s2[p], erow[p]), zero2);
            hd = hrow[p];
            dd[p] = d;
            dm[p] = d + pk_splat((p + 1) * e_ins - oe_ins);
            grun = pk_max(grun, dm[p]);
        }
        grun = grun - pk_splat(KHe);  // F leaving each half's last pair
        const int32_t c_lo = max((int32_t)grun.x, 0), c_hi = max((int32_t)grun.y, 0);
        const int32_t U = max(c_hi, c_lo - KHe) + g * Ke;
        const int32_t incl = L::scan_max_nonneg(U);
        const int32_t excl = L::shift_in(incl, -Ke);
        const int32_t f_in = excl - (g - 1) * Ke;
        pk16 f = (pk16){(short)f_in, (short)max(f_in - KHe, c_lo)};  // F + p*e_ins at pair p
        pk16 key2 = zero2;
#pragma unroll
        for (int p = 0; p < KH; ++p) {
            const pk16 h = pk_max(dd[p], p == 0 ? f : f - pk_splat(p * e_ins));
            key2 = pk_max(key2, pk_key(h, kk[p]));
            erow[p] = pk_max(erow[p] - e_del2, h - oe_del2);
            f = pk_max(f, dm[p]);
            hrow[p] = h;
        }
        const uint32_t kbest = (uint32_t)max((int32_t)key2.x, (int32_t)key2.y);  // >= 0
        const int32_t key = L::allmax((int32_t)(__umul24(kbest >> 6, 960u) + kbest + g64));
        const int32_t imax = key >> 10;
        if (!REV && __builtin_expect(g == 0, 1)) *rm = imax;  // under the exec mask, no branch
        if (!REV) rm += 64;  // row i+1's slot (a running pointer: one 64-bit add a row)
        bool done = false;
        if (imax > gmax) {
            gmax = imax;
            if (REV) {
                if (gmax >= endsc) {
                    // first row reaching the forward score: the alignment starts rte/rqe back
                    if (g == 0 && gmax == endsc) {
                        const int32_t rqe = pk_key_column<GL>(key, K);
                        out[idx].tb = te - i;
                        out[idx].qb = qe - rqe;
                    }
                    done = true;
                }
            } else {
                te = i;
                qe = pk_key_column<GL>(key, K);
            }
        }
        ++i;
        if (i >= tlen) done = true;
        if (done) {
            if (!REV && g == 0) {
                bwg_ksw_result r;
                r.score = gmax;
                r.qb = -1;
                r.qe = te >= 0 ? qe : -1;
                r.tb = -1;
                r.te = te;
                r.score2 = -1;
                r.te2 = -1;
                out[idx] = r;
            }
            active = false;
        }
    }
}

// score2/te2 from the recorded row maxima (crates/bwa-extend/src/sw.rs:806-851): replay
// push_row's merge rule and keep the best closed entry outside the exclusion window around te.
__global__ __launch_bounds__(256) void ksw_tracker_kernel(const KswJob *__restrict__ jobs,
                                                          const uint32_t *__restrict__ list,
                                                          uint32_t count, KswScoring sc,
                                                          const int32_t *__restrict__ rowmax,
                                                          bwg_ksw_result *__restrict__ out) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const uint32_t idx = list[k];
    const KswJob jb = jobs[idx];
    bwg_ksw_result r = out[idx];
    const int32_t gmax = r.score, te = r.te;
    int32_t s2 = -1, t2 = -1;
    if (gmax > 0 && jb.qlen > 0) {
        const int32_t *rm = rowmax + row_slot(jb);
        const int32_t half = (gmax + sc.max_sc - 1) / sc.max_sc;
        const int32_t lo = te - half, hi = te + half;
        int32_t cs = 0, cr = -1;
        bool open = false;
        const int32_t tlen = (int32_t)jb.tlen;
        // eight independent loads in flight per step; the merge rule then runs over them in order
        constexpr int kAhead = 8;
        for (int32_t i0 = 0; i0 < tlen; i0 += kAhead) {
            int32_t vv[kAhead];
#pragma unroll
            for (int u = 0; u < kAhead; ++u) vv[u] = i0 + u < tlen ? rm[64 * (i0 + u)] : 0;
#pragma unroll
            for (int u = 0; u < kAhead; ++u) {
                const int32_t i = i0 + u, v = vv[u];
                if (i >= tlen) break;
                if (v < sc.minsc) continue;
                if (open && cr + 1 == i) {
                    if (cs < v) {
                        cs = v;
                        cr = i;
                    }
                    continue;
                }
                if (open && (cr < lo || cr > hi) && cs > s2) {
                    s2 = cs;
                    t2 = cr;
                }
                cs = v;
                cr = i;
                open = true;
            }
        }
        if (open && (cr < lo || cr > hi) && cs > s2) {
            s2 = cs;
            t2 = cr;
        }
    }
    out[idx].score2 = s2;
    out[idx].te2 = t2;
}

template <int K, bool REV>
int occ() {
    static std::atomic<int> cache{0};
    int v = cache.load(std::memory_order_relaxed);
    if (v > 0) return v;
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, ksw_group_kernel<K, REV>, 256, 0) != hipSuccess)
        blocks = 1;
    v = blocks > 0 ? blocks : 1;
    cache.store(v, std::memory_order_relaxed);
    return v;
}

template <int KH, bool REV, int GL>
int occ_pk() {
    static std::atomic<int> cache{0};
    int v = cache.load(std::memory_order_relaxed);
    if (v > 0) return v;
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, ksw_pk_kernel<KH, REV, GL>, 256, 0) != hipSuccess)
        blocks = 1;
    v = blocks > 0 ? blocks : 1;
    cache.store(v, std::memory_order_relaxed);
    return v;
}

}  // namespace

int ksw_pk_blocks_per_cu(int param, bool rev) {
    const int K = ksw_param_k(param);
    if (ksw_param_lanes(param) == 8) {
        switch (K) {
#define BWG_K(k) case k: return rev ? occ_pk<k, true, 8>() : occ_pk<k, false, 8>();
            BWG_KSW_GROUP_KS
#undef BWG_K
        default: return 1;
        }
    }
    switch (K) {
#define BWG_K(k) case k: return rev ? occ_pk<k / 2, true, 16>() : occ_pk<k / 2, false, 16>();
        BWG_KSW_GROUP_KS
#undef BWG_K
    default: return 1;
    }
}

// param: the 16-lane K, or ksw_param(K, 8) for 8-lane groups with 2K columns (K pairs) per lane.
hipError_t launch_ksw_pk(hipStream_t st, int param, bool rev, const uint8_t *arena, const uint8_t *ref,
                         const KswJob *jobs, const uint32_t *order, uint32_t count, uint32_t *head,
                         const KswScoring &sc, bwg_ksw_result *out, int32_t *rowmax, int grid) {
    const int K = ksw_param_k(param);
#define BWG_LAUNCH(KH, GL)                                                                                    \
    do {                                                                                                      \
        if (rev)                                                                                              \
            hipLaunchKernelGGL((ksw_pk_kernel<KH, true, GL>), dim3(grid), dim3(256), 0, st, arena, ref, jobs,  \
                               order, count, head, sc, out, rowmax);                                          \
        else                                                                                                  \
            hipLaunchKernelGGL((ksw_pk_kernel<KH, false, GL>), dim3(grid), dim3(256), 0, st, arena, ref, jobs, \
                               order, count, head, sc, out, rowmax);                                          \
    } while (0)
    if (ksw_param_lanes(param) == 8) {
        switch (K) {
#define BWG_K(k)              \
    case k:                   \
        BWG_LAUNCH(k, 8);     \
        return hipGetLastError();
            BWG_KSW_GROUP_KS
#undef BWG_K
        default: return hipErrorInvalidValue;
        }
    }
    switch (K) {
#define BWG_K(k)              \
    case k:                   \
        BWG_LAUNCH(k / 2, 16); \
        return hipGetLastError();
        BWG_KSW_GROUP_KS
#undef BWG_K
    default: return hipErrorInvalidValue;
    }
#undef BWG_LAUNCH
}

int ksw_group_blocks_per_cu(int K, bool rev) {
    switch (K) {
#define BWG_K(k) case k: return rev ? occ<k, true>() : occ<k, false>();
        BWG_KSW_GROUP_KS
#undef BWG_K
    default: return 1;
    }
}

hipError_t launch_ksw_group(hipStream_t st, int K, bool rev, const uint8_t *arena, const uint8_t *ref, const KswJob *jobs,
                            const uint32_t *order, uint32_t count, uint32_t *head, const KswScoring &sc,
                            bwg_ksw_result *out, int32_t *rowmax, int grid) {
    switch (K) {
#define BWG_K(k)                                                                                          \
    case k:                                                                                               \
        if (rev)                                                                                          \
            hipLaunchKernelGGL((ksw_group_kernel<k, true>), dim3(grid), dim3(256), 0, st, arena, ref, jobs, \
                               order, count, head, sc, out, rowmax);                                      \
        else                                                                                              \
            hipLaunchKernelGGL((ksw_group_kernel<k, false>), dim3(grid), dim3(256), 0, st, arena, ref, jobs, \
                               order, count, head, sc, out, rowmax);                                      \
        return hipGetLastError();
        BWG_KSW_GROUP_KS
#undef BWG_K
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_ksw_tracker(hipStream_t st, const KswJob *jobs, const uint32_t *list, uint32_t count,
                              const KswScoring &sc, const int32_t *rowmax, bwg_ksw_result *out) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(ksw_tracker_kernel, dim3((count + 255) / 256), dim3(256), 0, st, jobs, list, count,
                       sc, rowmax, out);
    return hipGetLastError();
}

}  // namespace bwg
