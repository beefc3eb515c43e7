// Batched form of the banded extension (ksw_extend2 semantics, crates/bwa-extend/src/sw.rs:203-493).
//
// Two persistent kernels, each draining a device work queue of length-binned jobs:
//
//  * ext_group_kernel<K>: one job per 16-lane DPP row ("group"), four jobs per wavefront.  Row i
//    of the DP is evaluated in parallel over the live window [beg, end): the lanes own blocks of kr
//    consecutive columns aligned on the window's right end (window-relative blocking), where
//    kr <= K is chosen per row as the smallest odd width that covers the widest window among the
//    wavefront's four groups (odd keeps the 16 lanes' LDS accesses in distinct banks; K bounds
//    the registers).  The eh arrays of the reference live in LDS over absolute columns 0..qlen, so
//    columns past the window keep their stale values exactly as in the reference; the slots below
//    the window are kept at zero, so the first block's columns left of the window compute zeros
//    and every column store is unconditional.  M and E
//    depend only on the previous row; F is the max-plus recurrence
//    F(j+1) = max(F(j) - e_ins, max(M(j) - oe_ins, 0)), evaluated in two passes: each lane runs it
//    over its block from F_in = 0, the blocks are combined by an exclusive max-scan across the row
//    (F(l) = max_{m<l} c_m - (l-1-m)*kr*e_ins), and a second pass applies each lane's incoming F.
//    Row maximum with the later-column tie rule is a max over keys (h << 13 | j).  The row
//    epilogue (to-end score, z-drop, band tightening) is evaluated redundantly by the 16 lanes.
//    The template's GL = 8 runs the same row with 8-lane groups, two per DPP row interleaved by
//    lane parity (GroupLanes<8>), so eight jobs share a wavefront and its per-row bookkeeping; its
//    K (up to 31) is the columns per lane for the same windows.  It is an autotuner candidate.
//
//  * ext_thread_kernel<QCAP>: one job per lane for short queries (qlen <= QCAP), the scalar
//    recurrence with the eh arrays in LDS laid out [column][lane].
//
// Both deal most of their queue's slots round-robin to the grid (GroupSlots; the thread kernel in
// 64-slot chunks per wavefront) and take the grid's last slots from the queue's counter.
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
    int32_t qlen, tlast, h0, w;  // tlast = tlen - 1
    uint32_t out_idx;
    // row state
    int32_t i, beg, end, best, best_i, best_j, gbest_i, gscore, max_off;
    int32_t tcur;  // target code of row i
    uint32_t h1c;  // max(h0 - (o_del + e_del*(i+1)), 0), decremented with saturation row by row
};

// The row end of a group without a job: its window (end - beg) is negative whatever beg is, so
// the wavefront's widest window ignores it.
constexpr int32_t kIdleEnd = INT32_MIN / 2;

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

// LDS bank layout.  The LDS serves one half-wavefront (32 lanes) per cycle from 32 banks of 4
// bytes, so the NG = 32 / GL groups of a half-wavefront share each access.  Their per-column words
// are interleaved: column j of the group with index q = gid % NG within its half-wavefront sits at
// word NG*j + q of the half-wavefront's region.  A group's lanes read columns kr apart (kr odd),
// words NG*kr apart, which is NG times a permutation of 0..GL-1 modulo 32: the GL lanes hit GL
// distinct banks, all congruent to q modulo NG, so the groups of a half-wavefront never share a
// bank whatever their windows.
template <int GL> constexpr int groups_per_half() { return 32 / GL; }

// The column state of one group, the reference's eh arrays over absolute columns 0..qcap+1, below
// which ext_pad(K) zero columns pad the left end of the row blocks.  Wide: H and E as int32, each in its
// own interleaved array.  Narrow: one word per column, H in its low and E in its high 16 bits,
// exact whenever H and E stay below 65536; they never exceed h0 + qlen*max_sc, which the planner
// checks for narrow bins.  Narrow state halves the group's LDS, which is what limits the kernel's
// occupancy for wide windows.  `col0` is the group's word of column 0; column j is NG words further
// per column.
template <bool NARROW, int NG> struct EhView;
template <int NG> struct EhView<false, NG> {
    int32_t *hp, *ep;
    __device__ __forceinline__ EhView(int32_t *col0, int qcap, int pad)
        : hp(col0), ep(col0 + NG * (pad + qcap + 2)) {}
    __device__ __forceinline__ int2 load(int j) const { return make_int2(hp[NG * j], ep[NG * j]); }
    __device__ __forceinline__ int32_t load_h(int j) const { return hp[NG * j]; }
    __device__ __forceinline__ int32_t load_e(int j) const { return ep[NG * j]; }
    __device__ __forceinline__ void set_h(int j, int32_t v) const { hp[NG * j] = v; }
    __device__ __forceinline__ void set_e(int j, int32_t v) const { ep[NG * j] = v; }
    __device__ __forceinline__ void init(int j, int32_t h) const {
        hp[NG * j] = h;
        ep[NG * j] = 0;
    }
};
template <int NG> struct EhView<true, NG> {
    uint16_t *p;  // H of column j at p[2*NG*j], E at p[2*NG*j + 1]
    __device__ __forceinline__ EhView(int32_t *col0, int, int) : p((uint16_t *)col0) {}
    // The widened values are pinned where they are loaded: widening a 16-bit value in a later block
    // (where the column's use sits behind a branch) costs an AND there, where ds_read_u16 has
    // already zero-extended it.
    __device__ __forceinline__ int2 load(int j) const {
        int32_t h = p[2 * NG * j], e = p[2 * NG * j + 1];
        asm("" : "+v"(h));
        asm("" : "+v"(e));
        return make_int2(h, e);
    }
    __device__ __forceinline__ int32_t load_h(int j) const {
        int32_t h = p[2 * NG * j];
        asm("" : "+v"(h));
        return h;
    }
    __device__ __forceinline__ int32_t load_e(int j) const {
        int32_t e = p[2 * NG * j + 1];
        asm("" : "+v"(e));
        return e;
    }
    __device__ __forceinline__ void set_h(int j, int32_t v) const { p[2 * NG * j] = (uint16_t)v; }
    __device__ __forceinline__ void set_e(int j, int32_t v) const { p[2 * NG * j + 1] = (uint16_t)v; }
    __device__ __forceinline__ void init(int j, int32_t h) const {
        *(uint32_t *)(p + 2 * NG * j) = (uint32_t)(uint16_t)h;
    }
};

// Block LDS layout, in 4-byte words, for G = 256/GL groups in 8 half-wavefronts of NG groups:
//  * eh: G * eh_words(qcap, K) words; half-wavefront hw owns NG * eh_words words from
//    hw * NG * eh_words, interleaved as above, column j of group q at word NG*(ext_pad(K) + j) + q (the
//    wide form's E array follows its H array);
//  * score planes 0-3: G * qcap words; half-wavefront hw owns NG * qcap words from hw * NG * qcap,
//    word NG*j + q holding mat[c][query[j]] in byte c for target codes c = 0..3;
//  * score plane 4 (target N): G * qcap bytes; the half-wavefronts hw and hw + 1, hw + 2, hw + 3
//    (hw a multiple of 4) share NG * qcap words, word NG*j + q, byte hw & 3, so that every plane has
//    the same column stride of 4*NG bytes.
// A row touches columns from -ext_pad(K) up to end <= qlen: the eh pad covers the columns below 0,
// and the score bytes read there (multiplied by a zero H) come from the region before the plane's.
// Zero columns below column 0: a multiple of 16 and at least K + 15, so that a row block moved
// below the window keeps its column's residue modulo 16 (its LDS banks) and stays below column 0.
__host__ __device__ constexpr int ext_pad(int K) { return (K + 30) / 16 * 16; }
__host__ __device__ __forceinline__ int eh_words(int qcap, int K, bool narrow) {
    return (narrow ? 1 : 2) * (ext_pad(K) + qcap + 2);
}
__host__ __device__ __forceinline__ size_t block_lds_bytes(int GL, int qcap, int K, bool narrow) {
    const size_t groups = (size_t)(256 / GL);
    return groups * (4 * (size_t)eh_words(qcap, K, narrow) + 5 * (size_t)qcap) + 64;
}

// The cells of one DP row for one group, with the row's window [beg, end) spread over the group's
// lanes in blocks of KR columns aligned on its right end: lane g owns the columns base .. base +
// KR - 1, base = end - (GL - g)*KR (the group's KR*GL columns cover the window).  Every slot below
// beg holds zero (the caller keeps it so, and pads PAD >= KR columns below column 0), and on zero
// inputs a column computes zero H and E and passes F = 0 on.  So the first block in the window
// may reach left of it onto zero columns, and a lane whose block lies wholly left of the window
// computes and stores zeros over zero slots (a block that would start below the pad is moved up
// into it, keeping its residue modulo 16, so the lanes' blocks still start at distinct banks):
// no column load or store needs a predicate, and such a block hands F = 0 on.  Columns at or past end are never touched: they keep their stale values, which a later row
// reads when the band widens again, as the reference's do.  `srow` is the row's score plane,
// column j at byte 4*NG*j (NG = 32/GL groups per half-wavefront).  Returns the lane's row-maximum
// key (h << 13 | absolute column).  A column left of the window has h = 0, so it cannot win a row
// whose maximum is positive; when the window is empty every column is below beg and every h is 0,
// so the key is below 1 << 13 (negative where the columns are below 0).
template <int KR, int GL, int PAD, class EH>
__device__ __forceinline__ int32_t row_cells(const EH &eh, const int8_t *srow, int32_t end, int g, int32_t e_ins,
                                             int32_t oe_ins, int32_t e_del, int32_t oe_del) {
    using L = GroupLanes<GL>;
    static_assert(PAD % 16 == 0 && PAD >= KR + 15, "a moved block must keep its banks and stay below 0");
    // the block's first column; a block that would start below the pad (nat < -PAD) takes the
    // column of the same residue modulo 16 among the pad's lowest 16, -PAD + (nat & 15), with one
    // and-or (such a block lies left of the window; a block reaching into it starts above -KR)
    const int32_t nat = end - (GL - g) * KR;
    const int32_t base = max(nat, (nat & 15) | -PAD);
    constexpr int kSStride = 4 * groups_per_half<GL>();
    const int8_t *sb = srow + base * kSStride;
    // pass 1: the loads are independent, so they issue back to back (the score bytes first: the
    // column loads end in the 16-bit state's widening pins, which wait for them)
    int2 v[KR];
    int32_t s[KR];
#pragma unroll
    for (int k = 0; k < KR; ++k) s[k] = sb[k * kSStride];
#pragma unroll
    for (int k = 0; k < KR; ++k) v[k] = eh.load(base + k);
    // F from 0 at the block's first column, carried with its drift: G_k = F_k + k*e_ins, so
    // F_{k+1} = max(F_k - e_ins, M_k - oe_ins, 0) becomes G_{k+1} = max(G_k, M_k + (k+1)*e_ins -
    // oe_ins, (k+1)*e_ins), one add and one three-way max per column
    int32_t mm[KR], gg[KR];
    int32_t grun = 0;
#pragma unroll
    for (int k = 0; k < KR; ++k) {
        // M = H ? H + s : 0, as s*min(H, 1) + H (H >= 0): two operations, where a compare and
        // a select would take three (the min is opaque so that the compiler keeps this form)
        uint32_t z;
        asm("v_min_u32 %0, %1, 1" : "=v"(z) : "v"(v[k].x));
        const int32_t m = __mul24(s[k], (int32_t)z) + v[k].x;
        mm[k] = m;
        gg[k] = grun;
        grun = max(max(grun, m + ((k + 1) * e_ins - oe_ins)), (k + 1) * e_ins);
    }
    grun -= KR * e_ins;
    // F entering lane l's block: F(l) = max_{m<l} c_m - (l-1-m)*KR*e_ins over the blocks' exit
    // values c_m >= 0 (0 for a block of zero columns), as an exclusive max-scan of U_l = c_l +
    // (l+1)*KR*e_ins (>= 0, so the scan's zero fill is its identity), F(l) = scan(l) - l*KR*e_ins.
    // The first lane gets 0, and a term from a zero block is <= 0, which the F floor in gg absorbs.
    const int32_t Ke = KR * e_ins;
    const int32_t incl = L::scan_max_nonneg(grun + (g + 1) * Ke);
    const int32_t f_in = L::shift_in_zero(incl) - g * Ke;

    int32_t key = 0;
    int32_t hh[KR], ee[KR];
#pragma unroll
    for (int k = 0; k < KR; ++k) {
        const int32_t f = max(f_in, gg[k]) - k * e_ins;
        const int32_t h = max(max(mm[k], v[k].y), f);
        hh[k] = h;
        ee[k] = max(max(v[k].y - e_del, mm[k] - oe_del), 0);
        // lane-relative column; the lane's first column is added once below
        key = k == 0 ? h << kKeyShift : max(key, (h << kKeyShift) | k);
    }
#pragma unroll
    for (int k = 0; k < KR; ++k) {
        eh.set_e(base + k, ee[k]);      // E(i+1, j)
        eh.set_h(base + k + 1, hh[k]);  // H(i, j): the diagonal of column j+1 next row
    }
    return key + base;
}

// row_cells at the odd block width kr (lo <= kr <= hi, both odd), chosen by a binary search
// over the wavefront-uniform kr so that each row runs one straight-line body.
template <int LO, int HI, int GL, int PAD, class EH>
__device__ __forceinline__ int32_t row_cells_at(int kr, const EH &eh, const int8_t *srow, int32_t end, int g,
                                                int32_t e_ins, int32_t oe_ins, int32_t e_del, int32_t oe_del) {
    if constexpr (LO >= HI) {
        return row_cells<HI, GL, PAD>(eh, srow, end, g, e_ins, oe_ins, e_del, oe_del);
    } else {
        constexpr int MID = LO + ((HI - LO) / 4) * 2;  // odd, LO <= MID < HI
        if (kr <= MID) return row_cells_at<LO, MID, GL, PAD>(kr, eh, srow, end, g, e_ins, oe_ins, e_del, oe_del);
        return row_cells_at<MID + 2, HI, GL, PAD>(kr, eh, srow, end, g, e_ins, oe_ins, e_del, oe_del);
    }
}

// Groups of a 256-thread block: 16 of 16 lanes, or 32 of 8 lanes.
template <int GL>
constexpr int groups_per_block() {
    return 256 / GL;
}

// Waves per SIMD the register allocation aims for: the 16-lane K = 11 kernel with 16-bit state
// fits 96 registers (5 waves; its LDS at the class's query lengths allows 5-6 blocks a CU, one
// register spilled), where it would otherwise take 100 (4 waves).  The same push for K = 7 and
// K = 9 (6 waves at 80 registers) was slower.
template <int K, bool NARROW, int GL>
constexpr int ext_group_min_waves() {
    return (K == 11 && NARROW && GL == 16) ? 5 : 1;
}

template <int K, bool NARROW, int GL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ext_group_min_waves<K, NARROW, GL>())))
void ext_group_kernel(const uint8_t *__restrict__ arena,
                                                        const uint8_t *__restrict__ ref,
                                                        const ExtJob *__restrict__ jobs,
                                                        const uint32_t *__restrict__ order,
                                                        uint32_t count_static, const uint32_t *__restrict__ dcount,
                                                        uint32_t *__restrict__ head, ExtScoring sc, int qcap,
                                                        bwg_ext_result *__restrict__ out) {
    extern __shared__ int32_t lds[];
    const uint32_t count = dcount ? *dcount : count_static;  // a compacted queue's length is on the device
    // an empty queue (a retry round that retries nothing) ends the grid before any queue atomic
    if (count == 0) return;
    using L = GroupLanes<GL>;
    const int g = L::g();      // lane within the group
    const int gid = L::gid();  // group within the block
    // the group's place in the bank-interleaved block layout (see block_lds_bytes)
    constexpr int NG = groups_per_half<GL>();
    constexpr int G = groups_per_block<GL>();
    const int hw = gid / NG, q = gid % NG;
    int32_t *const lw = (int32_t *)lds;
    const int ehw = eh_words(qcap, K, NARROW);
    constexpr int PAD = ext_pad(K);
    const EhView<NARROW, NG> eh(lw + hw * NG * ehw + NG * PAD + q, qcap, PAD);
    int32_t *const plane03 = lw + G * ehw + hw * NG * qcap + q;  // word of column 0, byte c = code c
    int8_t *const plane4 = (int8_t *)(lw + G * ehw + G * qcap + (hw >> 2) * NG * qcap + q) + (hw & 3);

    const int32_t oe_del = sc.o_del + sc.e_del, oe_ins = sc.o_ins + sc.e_ins;
    const int32_t e_del = sc.e_del, e_ins = sc.e_ins;

    GroupJob J = {};
    bool active = false, exhausted = false;
    GroupSlotsT<GL> slots(count, true);
    for (;;) {
        // ---- fetch and initialise a job for idle groups ------------------------------------
        if (!active && !exhausted) {
            const uint32_t slot = slots.take(head);
            if (slot >= count) {
                exhausted = true;
                J.end = kIdleEnd;  // a done group takes a new job or gets here on the next pass
            } else {
                const ExtJob jb = jobs[order[slot]];
                const uint8_t *qry = arena + jb.q_off;
                J.t = target_view(arena, ref, jb.t_off, jb.tlen);
                J.qlen = (int32_t)jb.qlen;
                J.tlast = (int32_t)jb.tlen - 1;
                J.h0 = jb.h0;
                J.out_idx = order[slot];
                J.w = clamp_band(sc.w, J.qlen, sc);
                // row -1 (sw.rs:263-279): H[0] = h0, H[j] = max(h0 - oe_ins - (j-1)*e_ins, 0)
                // (and the zero pad below column 0, which the row blocks read left of the window)
                for (int j = g - PAD; j <= J.qlen + 1; j += GL) {
                    const int32_t h = j == 0 ? J.h0 : max(J.h0 - oe_ins - (j - 1) * e_ins, 0);
                    eh.init(j, j >= 0 && j <= J.qlen ? h : 0);
                }
                for (int j = g; j < J.qlen; j += GL) {
                    const int qc = qry[j];
                    plane03[NG * j] = (int32_t)((uint32_t)(uint8_t)sc.mat[qc] | (uint32_t)(uint8_t)sc.mat[5 + qc] << 8 |
                                                (uint32_t)(uint8_t)sc.mat[10 + qc] << 16 |
                                                (uint32_t)(uint8_t)sc.mat[15 + qc] << 24);
                    plane4[4 * NG * j] = sc.mat[20 + qc];
                }
                J.i = 0;
                J.beg = 0;
                J.end = J.qlen;
                J.best = J.h0;
                J.best_i = J.best_j = J.gbest_i = J.gscore = -1;
                J.max_off = 0;
                J.tcur = jb.tlen > 0 ? J.t[0] : 0;
                // H(i, -1) = max(h0 - (o_del + e_del*(i+1)), 0): its value at row 0, then minus
                // e_del a row, saturating at 0 (e_del >= 1, so it never rises again)
                J.h1c = (uint32_t)max((int32_t)((uint32_t)J.h0 - (uint32_t)sc.o_del - (uint32_t)e_del), 0);
                active = jb.tlen > 0;
                if (!active) {
                    J.end = kIdleEnd;
                    if (g == 0) {
                        bwg_ext_result r = {J.h0, 0, 0, 0, -1, 0};
                        out[J.out_idx] = r;
                    }
                }
            }
        }
        wave_lds_sync();
        // the wavefront exits once none of its four groups has work left
        if (__ballot(active || !exhausted) == 0) break;

        // ---- one DP row ------------------------------------------------------------------------
        const int32_t i = J.i;
        // an idle group's window is negative (kIdleEnd) or empty (no job yet)
        const int32_t beg = max(J.beg, i - J.w);
        const int32_t end = min(min(J.end, i + J.w + 1), J.qlen);
        // Block width of this row: the widest window among the wavefront's four groups, in an
        // odd number of columns per lane (at most K, since the bin's K covers the job's widest
        // possible window).  The band narrows as rows go by, and the wavefront's time per row
        // follows its widest group, so the row's columns are spread over every lane again.  A
        // wavefront whose groups all have empty windows takes the narrowest body (kr <= 1).
        const int32_t wmax = L::wave_max(end - beg);
        const int32_t kr = min(((wmax + GL - 1) / GL) | 1, K);
        if (!active) continue;
        // Every slot below the window's first column holds zero, which the row blocks rely on.
        // Tightening only moves beg past dead (zero) slots; the band moves it at most one column
        // a row (i - w <= J.beg + 1), past a slot that may be live and that no row reads again:
        // slot beg - 1 is cleared every row, which is that slot when the band moved beg and a zero
        // slot (or the pad) otherwise.  The group's lanes all store the same value to the same
        // address, so the edge stores need no exec mask.
        eh.init(beg - 1, 0);
        const int32_t h1_init = beg == 0 ? (int32_t)J.h1c : 0;
        const int8_t *srow = J.tcur < 4 ? (const int8_t *)plane03 + J.tcur : plane4;
        // This row's target base (requested a row ago) is consumed before the next one is
        // requested: the vector-memory counter is in order, so a wait placed after the new request
        // would wait for it too, a full memory round trip every row.  The next row's base is
        // loaded into the same variable, so no copy at the row's end waits for it.
        asm volatile("" ::"v"(J.tcur) : "memory");
        J.tcur = J.t[min(i + 1, J.tlast)];
        int32_t key = row_cells_at<1, K, GL, PAD>(kr, eh, srow, end, g, e_ins, oe_ins, e_del, oe_del);
        eh.set_h(beg, h1_init);  // H(i, beg-1)
        eh.set_e(end, 0);        // E at the sentinel column
        wave_lds_sync();
        key = L::allmax(key);  // h << 13 | absolute column
        const int32_t row_max = key >> kKeyShift;
        const int32_t mj = key & kKeyMask;
        // the edge slots: end holds H(i, end-1) (h1_init if the window is empty) and the E just
        // reset to 0; beg holds h1_init and E(i+1, beg)
        const int32_t h1_last = eh.load_h(end);
        const int32_t e_beg = eh.load_e(beg);

        const int32_t j_exit = beg < end ? end : beg;
        if (j_exit == J.qlen && J.gscore <= h1_last) {
            J.gbest_i = i;
            J.gscore = h1_last;
        }
        // a row whose maximum is 0 ends the extension (an empty window's key is below 1 << 13);
        // a greater maximum moves the best cell (selects, not a branch: it is taken on most rows);
        // a lower one may trigger z-drop.  |mj - i| as one sum of absolute differences: mj < 8192
        // and i < 65536 fit its 16-bit operands.
        bool done = key < (1 << kKeyShift);
        const bool improve = row_max > J.best;
        const int32_t off = (int32_t)__builtin_amdgcn_sad_u16((uint32_t)mj, (uint32_t)i, 0u);
        J.max_off = improve ? max(J.max_off, off) : J.max_off;
        J.best_i = improve ? i : J.best_i;
        J.best_j = improve ? mj : J.best_j;
        J.best = improve ? row_max : J.best;
        if (!done && !improve && sc.zdrop > 0 && J.best - row_max > sc.zdrop) {
            // the gap term is never negative, so a drop within zdrop cannot end the extension
            const int32_t di = i - J.best_i, dj = mj - J.best_j;
            const int32_t dd = di > dj ? di - dj : dj - di, ge = di > dj ? e_del : e_ins;
            const int32_t gap = sc.small_ext ? __mul24(dd, ge) : dd * ge;
            if (J.best - row_max - gap > sc.zdrop) done = true;
        }
        if (!done) {
            // band tightening (sw.rs:464-480): first live column in [beg, end), last live column
            // scanning down from end.  The edges are usually live, so test them first.  A row that
            // goes on has a positive cell, so beg < end here.
            int32_t nb = beg;
            if ((h1_init | e_beg) == 0) {
                for (;;) {
                    const int j = nb + g;
                    bool live = false;
                    if (j < end) {
                        const int2 x = eh.load(j);
                        live = (x.x | x.y) != 0;
                    }
                    const int32_t cand = L::allmin(live ? j : 0x7fffffff);
                    if (cand != 0x7fffffff) { nb = cand; break; }
                    nb += GL;
                    if (nb >= end) { nb = end; break; }
                }
            }
            int32_t ne = end;
            if (h1_last == 0) {
                for (;;) {
                    const int j = ne - g;
                    bool live = false;
                    if (j >= nb) {
                        const int2 x = eh.load(j);
                        live = (x.x | x.y) != 0;
                    }
                    const int32_t cand = L::allmax(live ? j : -1);
                    if (cand >= 0) { ne = cand; break; }
                    ne -= GL;
                    if (ne < nb) { ne = nb - 1; break; }
                }
            }
            J.beg = nb;
            J.end = ne + 2;  // the next row's window clamps it to qlen
            J.i = i + 1;
            J.h1c = __builtin_elementwise_sub_sat(J.h1c, (uint32_t)e_del);
            if (i >= J.tlast) done = true;
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
                                                        uint32_t count_static, const uint32_t *__restrict__ dcount,
                                                        uint32_t *__restrict__ head, ExtScoring sc,
                                                        bwg_ext_result *__restrict__ out) {
    const uint32_t count = dcount ? *dcount : count_static;
    if (count == 0) return;  // an empty compacted queue: no queue atomics, no table set-up
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
    // 64 consecutive slots per wavefront: dealt round-robin (wavefront w takes 64w, 64(w + waves),
    // ...) up to the grid's last round, which comes from the queue counter (see GroupSlots)
    const uint32_t stride = gridDim.x * 64u;
    const uint32_t static_end = count > stride ? (count - stride) & ~63u : 0u;
    uint32_t next = blockIdx.x * 64u;
    for (;;) {
        uint32_t slot = 0;
        if (next < static_end) {
            slot = next;
            next += stride;
        } else {
            if (lane == 0) slot = static_end + atomicAdd(head, 64u);
            slot = __shfl(slot, 0, 64);
        }
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
                // the next column's state and query code are loaded a column ahead: column j
                // stores only slot j, so the loads do not wait on the store and their latency
                // overlaps this column's arithmetic
                int32_t j = beg;
                W vn = j < end ? eh[j * 64 + lane] : W{};
                uint32_t cn = j < end ? codes[j * 64 + lane] : 0u;
                for (; j < end; ++j) {
                    const W v = vn;
                    const uint32_t qc = cn;
                    if (j + 1 < end) {
                        vn = eh[(j + 1) * 64 + lane];
                        cn = codes[(j + 1) * 64 + lane];
                    }
                    const int32_t hd = CS::h_of(v), ev = CS::e_of(v);
                    const int32_t m = hd != 0 ? hd + subst(trow, qc) : 0;
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

// ---- ungapped-diagonal fast path -------------------------------------------------------------
// The reference's batched extension first tries a closed-form walk along the main diagonal
// (crates/bwa-neon/src/batched.rs:602-697, ungapped_hit) and runs the banded DP only for the jobs it
// declines.  Two kernels:
//
//  * ext_ungapped_kernel walks the jobs in INDEX order, one job per 16-lane group, so that one
//    wavefront's four groups read four consecutive 32-byte descriptors and, for a batch staged in
//    job order, neighbouring stretches of the arena: every fetched line is used, where a walk in bin
//    order touches a fresh line for every descriptor, query and target it reads.  Lane g reads
//    columns g, g+16, ... of the query and of the target's diagonal only (16 consecutive bytes per
//    step, the first 128 columns loaded up front); each lane counts the mismatches and N bases of
//    its columns, the group sums them once, and a job that passes with mismatches rebuilds their
//    columns (< 128) as a 128-bit mask from ballots.  An accepted job's result is written
//    directly; every job's verdict goes to one byte of `declined`.
//  * ext_compact_kernel walks the order array's contiguous slot chunks and appends the declined
//    jobs, in slot order within each chunk, to their bins' compacted queues with one atomic per bin
//    (the counters sit kCountStride words apart, on separate cache lines).  Slot ranges
//    [seg_off[b], seg_off[b+1]) belong to bin b; bins with skip[b] set (the plain bin) are left
//    alone: that bin runs every one of its jobs from the order array and writes each result again,
//    the same value the walk wrote for the ones it accepted.
constexpr int kUngappedChunkMax = 2048;
constexpr int kUngappedBinsMax = 256;

// Bin of `slot`, searching forward from `bin` (slots only grow).
__device__ __forceinline__ int bin_of_slot(uint32_t slot, int bin, const uint32_t *seg_off, int n_bins) {
    while (bin + 1 < n_bins && slot >= seg_off[bin + 1]) ++bin;
    return bin;
}

__global__ __launch_bounds__(256) void ext_ungapped_kernel(const uint8_t *__restrict__ arena,
                                                           const uint8_t *__restrict__ ref,
                                                           const ExtJob *__restrict__ jobs, uint32_t n_jobs,
                                                           int32_t a, int32_t b, int32_t xmax, int32_t zdrop,
                                                           bwg_ext_result *__restrict__ out,
                                                           uint8_t *__restrict__ declined,
                                                           uint32_t *__restrict__ clear, uint32_t n_clear) {
    // the launch sequence's queue heads and counters, zeroed here instead of by a kernel of their
    // own ahead of this one (nothing reads them before the compaction that follows)
    if (blockIdx.x == 0)
        for (uint32_t k = threadIdx.x; k < n_clear; k += blockDim.x) clear[k] = 0u;
    const int lane = threadIdx.x & 63;
    const int g = lane & 15;
    const uint32_t stride = gridDim.x * 16u;
    static_assert(sizeof(ExtJob) == 32, "a descriptor is two 16-byte words");
    // One job's walk; the loop below alternates two descriptor buffers, so that the next job's
    // descriptor is in flight while this job's bytes load and no register copies move it between
    // iterations.
    auto walk = [&](const ExtJob &jb, const uint32_t idx) {
        const uint8_t *q = arena + jb.q_off;
        const TargetView t = target_view(arena, ref, jb.t_off, jb.tlen);
        const int32_t qlen = (int32_t)jb.qlen, tlen = (int32_t)jb.tlen, h0 = jb.h0;
        bool bail = !(qlen > 0 && h0 > 0 && tlen >= qlen);
        uint32_t qv[8], tv[8];
        // the target's diagonal bytes g, g + 16, ... through a running pointer, one add a step
        const uint8_t *tp = t.p + (ptrdiff_t)g * t.step;
        const ptrdiff_t t16 = (ptrdiff_t)16 * t.step;
#pragma unroll
        for (int st = 0; st < 8; ++st) {  // independent loads of the first 128 columns
            const int32_t c = 16 * st + g;
            const bool in = !bail && c < qlen;
            qv[st] = in ? q[c] : 0u;
            tv[st] = in ? *tp : 0u;
            tp += t16;
        }
        // Mismatches and N bases among the first 128 columns: each lane counts its own columns
        // (the zeros past qlen match), the group sums the counts and looks for a code >= 4 once.
        // Both tests are on totals, so the verdict is the one a column-by-column walk reaches.
        // The mismatch columns themselves are needed only by a job that has some and passes.
        uint32_t lmis = 0, lcodes = 0;
#pragma unroll
        for (int st = 0; st < 8; ++st) {
            lmis += qv[st] != tv[st] ? 1u : 0u;
            lcodes |= qv[st] | tv[st];
        }
        const int32_t nmis = row_allsum((int32_t)lmis);
        const uint64_t bn = __ballot(lcodes >= 4u);
        if (((uint32_t)(bn >> (lane & ~15)) & 0xffffu) || nmis > xmax) bail = true;
        uint32_t bits[4] = {0u, 0u, 0u, 0u};  // mismatch columns < 128 (bail and nmis are the group's)
        if (!bail && nmis > 0) {
#pragma unroll
            for (int st = 0; st < 8; ++st) {
                const uint64_t bm = __ballot(qv[st] != tv[st]);
                bits[st >> 1] |= ((uint32_t)(bm >> (lane & ~15)) & 0xffffu) << (16 * (st & 1));
            }
        }
        // past column 128 only a mismatch-free walk can be accepted
        for (int32_t c0 = 128; c0 < qlen && !bail; c0 += 16) {
            const int32_t c = c0 + g;
            const bool in = c < qlen;
            const uint32_t qb = in ? q[c] : 0u, tb = in ? t[c] : 0u;
            const uint64_t bx = __ballot(in && (qb >= 4 || tb >= 4 || qb != tb));
            if ((uint32_t)(bx >> (lane & ~15)) & 0xffffu) bail = true;
        }
        bwg_ext_result r;
        if (!bail && nmis == 0) {
            // a mismatch-free walk only climbs: local and to-end alignments both take the query
            r.score = h0 + qlen * a;
            r.qle = r.tle = r.gtle = qlen;
            r.gscore = r.score;
            r.max_off = 0;
        } else if (!bail) {
            // the walk in runs between mismatches: a run of matches only climbs, so its last
            // column is its best and z-drop can only fire at a mismatch; a dead diagonal declines.
            // With mismatches the walk covers at most 128 columns.
            bail = qlen > 128;
            int32_t running = h0, best = h0, best_end = 0, prev = 0;
            for (int w = 0; w < 4 && !bail; ++w) {
                uint32_t m = bits[w];
                while (m && !bail) {
                    const int32_t p = 32 * w + __ffs(m) - 1;
                    m &= m - 1;
                    running += (p - prev) * a;
                    if (p > prev && running > best) {
                        best = running;
                        best_end = p;
                    }
                    running -= b;
                    if (running <= 0) bail = true;
                    else if (zdrop > 0 && best - running > zdrop) bail = true;
                    prev = p + 1;
                }
            }
            if (!bail) {
                running += (qlen - prev) * a;
                if (qlen > prev && running > best) {
                    best = running;
                    best_end = qlen;
                }
                r.score = best;
                r.qle = r.tle = best_end;
                r.gtle = qlen;
                r.gscore = running;
                r.max_off = 0;
            }
        }
        if (g == 0) {
            if (!bail) out[idx] = r;
            declined[idx] = bail ? 1 : 0;
        }
    };
    uint32_t idx = blockIdx.x * 16u + (threadIdx.x >> 4);
    ExtJob ja{}, jb{};
    if (idx < n_jobs) ja = jobs[idx];
    // (the steps compare n_jobs - idx with the stride: no 32-bit index wraps)
    while (idx < n_jobs) {
        const bool more_a = n_jobs - idx > stride;
        if (more_a) jb = jobs[idx + stride];
        walk(ja, idx);
        if (!more_a) break;
        idx += stride;
        const bool more_b = n_jobs - idx > stride;
        if (more_b) ja = jobs[idx + stride];
        walk(jb, idx);
        if (!more_b) break;
        idx += stride;
    }
}

__global__ __launch_bounds__(256) void ext_compact_kernel(const uint32_t *__restrict__ order, uint32_t n_slots,
                                                          uint32_t chunk, const uint32_t *__restrict__ seg_off,
                                                          const uint32_t *__restrict__ skip, int n_bins,
                                                          const uint8_t *__restrict__ declined,
                                                          uint32_t *__restrict__ corder,
                                                          uint32_t *__restrict__ ccount) {
    __shared__ uint32_t s_bincnt[kUngappedBinsMax];  // declined jobs per bin (bins from the chunk's first)
    __shared__ uint32_t s_binpos[kUngappedBinsMax];  // their first compacted-queue position
    __shared__ uint32_t s_seg[kUngappedBinsMax];   // the bins' first slots and skip flags, read
    __shared__ uint32_t s_skip[kUngappedBinsMax];  // from LDS by every bin search below
    __shared__ uint32_t s_wave[4];
    const int lane = threadIdx.x & 63;
    const uint32_t lo = blockIdx.x * chunk;
    const uint32_t hi = min(n_slots, lo + chunk);
    const uint32_t n = hi > lo ? hi - lo : 0;
    for (int k = threadIdx.x; k < kUngappedBinsMax; k += blockDim.x) {
        s_bincnt[k] = 0;
        if (k < n_bins) {
            s_seg[k] = seg_off[k];
            s_skip[k] = skip[k];
        }
    }
    __syncthreads();
    const int bin0 = bin_of_slot(lo, 0, s_seg, n_bins);
    // thread t owns the slots [t*per, t*per + per) of the chunk: at most kPer of them
    constexpr uint32_t kPer = (kUngappedChunkMax + 255) / 256;
    const uint32_t per = (n + blockDim.x - 1) / blockDim.x;
    const uint32_t t0 = min(n, threadIdx.x * per), t1 = min(n, t0 + per);
    uint32_t idx[kPer];
    uint32_t flags = 0, cnt = 0;
    int bb = bin_of_slot(lo + t0, bin0, s_seg, n_bins);
    const int first_bb = bb;
#pragma unroll
    for (uint32_t u = 0; u < kPer; ++u) {
        const uint32_t k = t0 + u;
        idx[u] = 0;
        if (k < t1) {
            bb = bin_of_slot(lo + k, bb, s_seg, n_bins);
            idx[u] = order[lo + k];
            if (!s_skip[bb] && declined[idx[u]]) {
                flags |= 1u << u;
                ++cnt;
                atomicAdd(&s_bincnt[bb - bin0], 1u);
            }
        }
    }
    __syncthreads();
    // one reservation per bin of the chunk
    const int nb = n ? bin_of_slot(hi - 1, bin0, s_seg, n_bins) - bin0 + 1 : 0;
    for (int k = threadIdx.x; k < nb; k += blockDim.x) {
        const uint32_t c = s_bincnt[k];
        s_binpos[k] = c ? s_seg[bin0 + k] + atomicAdd(&ccount[kCountStride * (bin0 + k)], c) : 0u;
    }
    // exclusive scan of the declined counts in slot order
    uint32_t incl = cnt;  // inclusive scan over the wavefront
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(incl, d, 64);
        if (lane >= d) incl += y;
    }
    if (lane == 63) s_wave[threadIdx.x >> 6] = incl;
    __syncthreads();
    uint32_t before = incl - cnt;  // declined slots before t0
    for (int w = 0; w < (int)(threadIdx.x >> 6); ++w) before += s_wave[w];
    // rank within its bin = declined slots before it in the chunk minus those of the earlier bins
    if (cnt) {
        bb = first_bb;
        uint32_t bin_first = 0;  // declined slots of the chunk's bins before bb
        for (int k = 0; k < bb - bin0; ++k) bin_first += s_bincnt[k];
#pragma unroll
        for (uint32_t u = 0; u < kPer; ++u) {
            const uint32_t k = t0 + u;
            if (k < t1) {
                const int nbb = bin_of_slot(lo + k, bb, s_seg, n_bins);
                for (; bb < nbb; ++bb) bin_first += s_bincnt[bb - bin0];
                if (flags >> u & 1u) {
                    corder[s_binpos[bb - bin0] + (before - bin_first)] = idx[u];
                    ++before;
                }
            }
        }
    }
}

// ---- band-retry acceptance ------------------------------------------------------------------
// One thread per slot of the next round's order array.  The acceptance test is the reference's
// (crates/bwa-mem/src/across.rs:1444-1445; lib.rs:360): the score did not change from the previous
// band, or the best cell's diagonal offset stayed under (w>>1) + (w>>2).  Rejected jobs are appended
// to their bins' compacted queues with one atomic per bin and wavefront: slots are sorted by bin, so
// a wavefront's rejected lanes fall in a few bins, and the leader of each reserves their slots.
__global__ __launch_bounds__(256) void ext_retry_filter_kernel(
    const uint32_t *__restrict__ order, uint32_t n_slots, const uint32_t *__restrict__ seg_off,
    const uint32_t *__restrict__ skip, int n_bins, const bwg_ext_result *__restrict__ out,
    const int32_t *__restrict__ prev_in, int32_t *__restrict__ prev_out, uint8_t *__restrict__ active, int first,
    int32_t w_prev, int32_t w_next, int32_t *__restrict__ w_out, uint32_t *__restrict__ corder,
    uint32_t *__restrict__ ccount) {
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    bool queue = false;
    uint32_t idx = 0;
    int bin = 0;
    if (slot < n_slots) {
        while (bin + 1 < n_bins && slot >= seg_off[bin + 1]) ++bin;
        idx = order[slot];
        if (first || active[idx]) {
            const bwg_ext_result r = out[idx];
            const int32_t prev = first ? prev_in[idx] : prev_out[idx];
            const bool accept = r.score == prev || r.max_off < (w_prev >> 1) + (w_prev >> 2);
            active[idx] = accept ? 0 : 1;
            if (!accept) {
                prev_out[idx] = r.score;
                w_out[idx] = w_next;
                queue = !skip[bin];
            } else if (first) {
                w_out[idx] = w_prev;
            }
        }
    }
    uint64_t pending = __ballot(queue);
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const int lb = __shfl(bin, leader, 64);
        const uint64_t mine = __ballot(queue && bin == lb);
        uint32_t pos = 0;
        if (lane == leader) pos = atomicAdd(&ccount[kCountStride * lb], (uint32_t)__popcll(mine));
        pos = __shfl(pos, leader, 64);
        if (queue && bin == lb)
            corder[seg_off[lb] + pos + (uint32_t)__popcll(mine & ((1ull << lane) - 1ull))] = idx;
        pending &= ~mine;
    }
}

template <int K, bool NARROW, int GL>
hipError_t launch_group(hipStream_t st, const uint8_t *arena, const uint8_t *ref, const ExtJob *jobs, const uint32_t *order,
                        uint32_t count, const uint32_t *dcount, uint32_t *head, const ExtScoring &sc, int qcap,
                        bwg_ext_result *out, int grid) {
    const size_t lds = block_lds_bytes(GL, qcap, K, NARROW);
    hipLaunchKernelGGL((ext_group_kernel<K, NARROW, GL>), dim3(grid), dim3(256), lds, st, arena, ref, jobs, order,
                       count, dcount, head, sc, qcap, out);
    return hipGetLastError();
}

// Occupancy queries are cached: they depend only on the kernel and its LDS size.
template <int K, bool NARROW, int GL>
int group_occupancy(int qcap) {
    static std::atomic<int> cache[130];
    const int slot = std::min(qcap / 32, 129);
    int v = cache[slot].load(std::memory_order_relaxed);
    if (v > 0) return v;
    int blocks = 0;
    const size_t lds = block_lds_bytes(GL, qcap, K, NARROW);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, ext_group_kernel<K, NARROW, GL>, 256, lds) != hipSuccess)
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

size_t ext_group_lds_bytes(int K, int qcap, bool narrow) { return block_lds_bytes(16, qcap, K, narrow); }
size_t ext_group8_lds_bytes(int K, int qcap, bool narrow) { return block_lds_bytes(8, qcap, K, narrow); }

int ext_group_blocks_per_cu(int K, int qcap, bool narrow) {
    switch (K) {
#define BWG_K(k) case k: return narrow ? group_occupancy<k, true, 16>(qcap) : group_occupancy<k, false, 16>(qcap);
        BWG_EXT_GROUP_KS
#undef BWG_K
    default: return 1;
    }
}

int ext_group8_blocks_per_cu(int K, int qcap, bool narrow) {
    switch (K) {
#define BWG_K(k) case k: return narrow ? group_occupancy<k, true, 8>(qcap) : group_occupancy<k, false, 8>(qcap);
        BWG_EXT_GROUP8_KS
#undef BWG_K
    default: return 1;
    }
}

hipError_t launch_ext_group8(hipStream_t st, int K, bool narrow, const uint8_t *arena, const uint8_t *ref,
                             const ExtJob *jobs, const uint32_t *order, uint32_t count, const uint32_t *dcount,
                             uint32_t *head, const ExtScoring &sc, int qcap, bwg_ext_result *out, int grid) {
    switch (K) {
#define BWG_K(k)                                                                                                   \
    case k:                                                                                                        \
        return narrow ? launch_group<k, true, 8>(st, arena, ref, jobs, order, count, dcount, head, sc, qcap, out, grid) \
                      : launch_group<k, false, 8>(st, arena, ref, jobs, order, count, dcount, head, sc, qcap, out, grid);
        BWG_EXT_GROUP8_KS
#undef BWG_K
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_ext_group(hipStream_t st, int K, bool narrow, const uint8_t *arena, const uint8_t *ref,
                            const ExtJob *jobs, const uint32_t *order, uint32_t count, const uint32_t *dcount,
                            uint32_t *head, const ExtScoring &sc, int qcap, bwg_ext_result *out, int grid) {
    switch (K) {
#define BWG_K(k)                                                                                              \
    case k:                                                                                                   \
        return narrow ? launch_group<k, true, 16>(st, arena, ref, jobs, order, count, dcount, head, sc, qcap, out, grid) \
                      : launch_group<k, false, 16>(st, arena, ref, jobs, order, count, dcount, head, sc, qcap, out, grid);
        BWG_EXT_GROUP_KS
#undef BWG_K
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_ext_retry_filter(hipStream_t st, const uint32_t *order, uint32_t n_slots, const uint32_t *seg_off,
                                   const uint32_t *skip, int n_bins, const bwg_ext_result *out,
                                   const int32_t *prev_in, int32_t *prev_out, uint8_t *active, bool first,
                                   int32_t w_prev, int32_t w_next, int32_t *w_out, uint32_t *corder,
                                   uint32_t *ccount) {
    if (n_slots == 0) return hipSuccess;
    hipLaunchKernelGGL(ext_retry_filter_kernel, dim3((n_slots + 255) / 256), dim3(256), 0, st, order, n_slots,
                       seg_off, skip, n_bins, out, prev_in, prev_out, active, first ? 1 : 0, w_prev, w_next, w_out,
                       corder, ccount);
    return hipGetLastError();
}

hipError_t launch_ext_ungapped(hipStream_t st, const uint8_t *arena, const uint8_t *ref, const ExtJob *jobs,
                               const uint32_t *order, uint32_t n_slots, const uint32_t *seg_off,
                               const uint32_t *skip, int n_bins, uint8_t *declined, uint32_t *corder,
                               uint32_t *ccount, const ExtScoring &sc, int32_t xmax, bwg_ext_result *out,
                               int min_blocks, uint32_t *clear, uint32_t n_clear) {
    if (n_slots == 0) return n_clear ? launch_zero_u32(st, clear, n_clear) : hipSuccess;
    if (n_bins > 255) return hipErrorInvalidValue;
    // the walk: a grid-stride loop of 16 jobs per block and step
    const uint32_t walk_blocks = std::min<uint32_t>((uint32_t)std::max(1, min_blocks), (n_slots + 15) / 16);
    hipLaunchKernelGGL(ext_ungapped_kernel, dim3(walk_blocks), dim3(256), 0, st, arena, ref, jobs, n_slots,
                       (int32_t)sc.mat[0], -(int32_t)sc.mat[1], xmax, sc.zdrop, out, declined, clear, n_clear);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // the compaction: contiguous chunks of at most kUngappedChunkMax slots, about min_blocks / 4 of them
    uint32_t blocks = (uint32_t)std::max(1, min_blocks / 4);
    uint32_t chunk = (n_slots + blocks - 1) / blocks;
    chunk = std::min<uint32_t>((chunk + 15) / 16 * 16, kUngappedChunkMax);
    blocks = (n_slots + chunk - 1) / chunk;
    hipLaunchKernelGGL(ext_compact_kernel, dim3(blocks), dim3(256), 0, st, order, n_slots, chunk, seg_off, skip,
                       n_bins, declined, corder, ccount);
    return hipGetLastError();
}

int ext_thread_blocks_per_cu(int qcap, bool narrow) {
    switch (qcap) {
    case 16: return narrow ? thread_occupancy<16, true>() : thread_occupancy<16, false>();
    case 32: return narrow ? thread_occupancy<32, true>() : thread_occupancy<32, false>();
    default: return 1;
    }
}

hipError_t launch_ext_thread(hipStream_t st, int qcap, bool narrow, const uint8_t *arena, const uint8_t *ref,
                             const ExtJob *jobs, const uint32_t *order, uint32_t count, const uint32_t *dcount,
                             uint32_t *head, const ExtScoring &sc, bwg_ext_result *out, int grid) {
#define BWG_T(Q, N)                                                                                        \
    hipLaunchKernelGGL((ext_thread_kernel<Q, N>), dim3(grid), dim3(64), 0, st, arena, ref, jobs, order, count, \
                       dcount, head, sc, out)
    if (qcap == 16 && narrow) BWG_T(16, true);
    else if (qcap == 16) BWG_T(16, false);
    else if (qcap == 32 && narrow) BWG_T(32, true);
    else if (qcap == 32) BWG_T(32, false);
    else return hipErrorInvalidValue;
#undef BWG_T
    return hipGetLastError();
}

}  // namespace bwg
