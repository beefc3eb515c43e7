// Launchers of the batched-form kernels.
#pragma once
#include "bwg_internal.h"

// Odd K values (window columns per lane) the group kernel is instantiated for; 16*K columns.
// Even K values (padded query columns per lane) the rescue group kernel is instantiated for.
#define BWG_KSW_GROUP_KS BWG_K(2) BWG_K(4) BWG_K(6) BWG_K(8) BWG_K(10) BWG_K(12) BWG_K(14) BWG_K(16)
#define BWG_EXT_GROUP_KS BWG_K(3) BWG_K(5) BWG_K(7) BWG_K(9) BWG_K(11) BWG_K(13) BWG_K(15) BWG_K(17) BWG_K(21) BWG_K(25) BWG_K(33)
// Columns per lane of the 8-lane group kernel: 2K + 1 for the 16-lane K up to 15.
#define BWG_EXT_GROUP8_KS BWG_K(7) BWG_K(11) BWG_K(15) BWG_K(19) BWG_K(23) BWG_K(27) BWG_K(31)

namespace bwg {

// Extension kernels.  The group kernel sizes its per-row block width itself (at most K columns per
// lane); `narrow` selects 16-bit column state in both kernels (valid when h0 + qlen*max_sc <
// 65536).
constexpr size_t kMaxBlockLds = 160u << 10;  // LDS a gfx950 workgroup may allocate
size_t ext_group_lds_bytes(int K, int qcap, bool narrow);
int ext_group_blocks_per_cu(int K, int qcap, bool narrow);
// `count` jobs of `order` (or *dcount of them when dcount is not null: a compacted queue).
hipError_t launch_ext_group(hipStream_t st, int K, bool narrow, const uint8_t *arena, const uint8_t *ref,
                            const ExtJob *jobs, const uint32_t *order, uint32_t count, const uint32_t *dcount,
                            uint32_t *head, const ExtScoring &sc, int qcap, bwg_ext_result *out, int grid);
// The same kernel with 8-lane groups, eight jobs per wavefront (K up to 31 columns per lane): half
// the per-row bookkeeping instructions per job row, and twice the block's LDS.
size_t ext_group8_lds_bytes(int K, int qcap, bool narrow);  // at most kMaxBlockLds to launch
int ext_group8_blocks_per_cu(int K, int qcap, bool narrow);
hipError_t launch_ext_group8(hipStream_t st, int K, bool narrow, const uint8_t *arena, const uint8_t *ref,
                             const ExtJob *jobs, const uint32_t *order, uint32_t count, const uint32_t *dcount,
                             uint32_t *head, const ExtScoring &sc, int qcap, bwg_ext_result *out, int grid);
int ext_thread_blocks_per_cu(int qcap, bool narrow);
hipError_t launch_ext_thread(hipStream_t st, int qcap, bool narrow, const uint8_t *arena, const uint8_t *ref,
                             const ExtJob *jobs, const uint32_t *order, uint32_t count, const uint32_t *dcount,
                             uint32_t *head, const ExtScoring &sc, bwg_ext_result *out, int grid);
// Ungapped-diagonal fast path (uniform DNA matrices with a > 0, b > 0 and xmax >= 0 only; at most
// 255 bins): a walk over all n_slots jobs in index order gives accepted jobs their result and every
// job its verdict in declined[job] (n_slots bytes), then a compaction over the order array appends
// the declined jobs of the bins without skip set to corder[seg_off[b] + ccount[kCountStride*b]++], in
// slot order within each block's chunk.  The order array is a permutation of the n_slots jobs.
constexpr int kCountStride = 32;  // compacted-queue counters one 128-byte line apart
// `clear` (n_clear words, the launch sequence's queue heads and counters) is zeroed by the walk's
// first block, ahead of the compaction and the bins that use it.
hipError_t launch_ext_ungapped(hipStream_t st, const uint8_t *arena, const uint8_t *ref, const ExtJob *jobs,
                               const uint32_t *order, uint32_t n_slots, const uint32_t *seg_off,
                               const uint32_t *skip, int n_bins, uint8_t *declined, uint32_t *corder,
                               uint32_t *ccount, const ExtScoring &sc, int32_t xmax, bwg_ext_result *out,
                               int min_blocks, uint32_t *clear = nullptr, uint32_t n_clear = 0);
hipError_t launch_zero_u32(hipStream_t st, uint32_t *p, uint32_t n);
hipError_t launch_fill_i32(hipStream_t st, int32_t *p, uint32_t n, int32_t v);
// Band-retry acceptance between two rounds of a one-sided extension batch (the reference's
// run_side, crates/bwa-mem/src/across.rs:1365-1460).  Over the next round's order array (bins
// delimited by seg_off): each job still active takes the acceptance test on its result at band
// w_prev against its previous score (prev_in in the first filter, prev_out after); an accepted job
// is deactivated, a rejected one keeps its score in prev_out, gets w_next in w_out and is appended to
// its bin's compacted queue unless the bin has skip set (those bins read `active` themselves).  The
// first filter treats every job as active and writes w_out for all of them.
hipError_t launch_ext_retry_filter(hipStream_t st, const uint32_t *order, uint32_t n_slots, const uint32_t *seg_off,
                                   const uint32_t *skip, int n_bins, const bwg_ext_result *out,
                                   const int32_t *prev_in, int32_t *prev_out, uint8_t *active, bool first,
                                   int32_t w_prev, int32_t w_next, int32_t *w_out, uint32_t *corder,
                                   uint32_t *ccount);

int ksw_group_blocks_per_cu(int K, bool rev);
hipError_t launch_ksw_group(hipStream_t st, int K, bool rev, const uint8_t *arena, const uint8_t *ref, const KswJob *jobs,
                            const uint32_t *order, uint32_t count, uint32_t *head, const KswScoring &sc,
                            bwg_ksw_result *out, int32_t *rowmax, int grid);
// Packed 16-bit form of the rescue pass (K columns per lane as K/2 pairs); exact for jobs with
// qlen * (largest matrix score) < kKswPkMaxScore under a scoring that ksw_pk_scoring_ok accepts:
// non-negative gap opens and open + extend below 2^14, so H*64 + 63 and every unfloored E and F
// stay inside int16; insertion extension below 4000, so F carried with its drift (up to eight
// extension penalties above F) does too.
constexpr int64_t kKswPkMaxScore = 500;
inline bool ksw_pk_scoring_ok(const KswScoring &sc) {
    return sc.o_del >= 0 && sc.o_ins >= 0 && sc.e_del >= 0 && sc.e_ins >= 0 && sc.o_del + sc.e_del < 16384 &&
           sc.o_ins + sc.e_ins < 16384 && sc.e_ins < 4000;
}
// A packed bin's param: the 16-lane K (2..16), or kKsw8Lanes + K for 8-lane groups with 2K
// columns (K pairs) per lane over the same padded query.  The 8-lane form carries F with up to 16
// extension penalties of drift, so it needs e_ins below 1000 (ksw_pk8_scoring_ok).
constexpr int kKsw8Lanes = 32;
inline int ksw_param(int K, int lanes) { return lanes == 8 ? kKsw8Lanes + K : K; }
inline int ksw_param_k(int param) { return param > kKsw8Lanes ? param - kKsw8Lanes : param; }
inline int ksw_param_lanes(int param) { return param > kKsw8Lanes ? 8 : 16; }
inline bool ksw_pk8_scoring_ok(const KswScoring &sc) { return ksw_pk_scoring_ok(sc) && sc.e_ins < 1000; }
int ksw_pk_blocks_per_cu(int param, bool rev);
hipError_t launch_ksw_pk(hipStream_t st, int param, bool rev, const uint8_t *arena, const uint8_t *ref,
                         const KswJob *jobs, const uint32_t *order, uint32_t count, uint32_t *head,
                         const KswScoring &sc, bwg_ksw_result *out, int32_t *rowmax, int grid);
hipError_t launch_ksw_tracker(hipStream_t st, const KswJob *jobs, const uint32_t *list, uint32_t count,
                              const KswScoring &sc, const int32_t *rowmax, bwg_ksw_result *out);

// Banded global alignment with traceback (global_batched.hip).  z_off is the job's first traceback
// dword (tlen rounded up to 8, times 16 * ((K+7)/8) of them), ops_off its first scratch word
// (qlen + tlen + 2 of them), k the group kernel's columns per lane for its bin.  `packed` selects
// the kernel form that keeps one byte per query column and the row's scores as 6-bit fields,
// valid when glb_scores_packable(sc) (every score in [-32, 31]).
struct GlobalScoring {
    int8_t mat[28];
    int32_t o_del, e_del, o_ins, e_ins;
};
struct GlobalJob {
    uint64_t q_off, t_off, z_off, ops_off;
    uint32_t qlen, tlen;
    int32_t w;
    uint32_t k;
};
bool glb_scores_packable(const GlobalScoring &sc);
size_t glb_group_lds_bytes(int K, int qcap, bool packed);
int glb_group_blocks_per_cu(int K, int qcap, bool packed);
hipError_t launch_glb_group(hipStream_t st, int K, bool packed, const uint8_t *arena, const uint8_t *ref,
                            const GlobalJob *jobs, const uint32_t *order, uint32_t count, uint32_t *head,
                            const GlobalScoring &sc, int qcap, uint32_t *z, int32_t *score, int grid);
// n_ops[k]: the job's CIGAR runs, back to front at ops + ops_off, or 0xffffffff when the walk left
// the stored band (possible only when no alignment path lies inside it).
hipError_t launch_glb_traceback(hipStream_t st, const GlobalJob *jobs, uint32_t n, const uint32_t *z, uint32_t *ops,
                                uint32_t *n_ops);
hipError_t launch_glb_cigar_gather(hipStream_t st, const GlobalJob *jobs, uint32_t n, const uint32_t *ops,
                                   const uint32_t *n_ops, const uint64_t *dst_off, uint32_t *dst);
// Long queries (k == 0 in their GlobalJob): one job per thread over `count` jobs of `order`, DP and
// traceback both, in a z region of 2*(qlen + 2) words plus tlen*min(qlen, 2w + 1) code bytes.
hipError_t launch_glb_plain(hipStream_t st, const uint8_t *arena, const uint8_t *ref, const GlobalJob *jobs,
                            const uint32_t *order, uint32_t count, const GlobalScoring &sc, uint32_t *z,
                            uint32_t *ops, uint32_t *n_ops, int32_t *score);
// The widest query the global group kernel keeps in LDS.
constexpr uint32_t kGlbGroupMaxQ = 512;

// Mate rescue with device-resident windows (rescue_windows.hip).  An anchor as the device reads it:
// the bwg_mate_anchor fields plus where its mate's reverse complement is written in the arena.
struct MateAnchor {
    int64_t rb;
    int32_t rid;
    uint32_t l_ms;
    uint64_t ms_off, rc_off, ma_begin, ma_end;
};
struct MateRescueParams {
    int64_t l_pac;
    int32_t low[4], high[4], failed[4];
    int32_t min_seed_len, n_contigs;
};
// ref2[l_pac + k] = complement(ref2[l_pac - 1 - k]) for k < l_pac.
hipError_t launch_reverse_half(hipStream_t st, uint8_t *ref2, int64_t l_pac);
hipError_t launch_mate_revcomp(hipStream_t st, uint8_t *arena, const MateAnchor *anchors, uint32_t n);
// Windows and descriptors of jobs 4k+r: jobs[] (indexed by job) and, for jobs in the plain bin,
// plain_jobs[plain_slot[job]] get the mate (or its reverse complement) as query and the clamped
// window of the two-strand reference as target (an empty target when the window is not valid).
hipError_t launch_matesw_windows(hipStream_t st, const MateAnchor *anchors, uint32_t n, const int64_t *ma_rb,
                                 const MateRescueParams &mp, const int64_t *contig_off, const int32_t *contig_len,
                                 KswJob *jobs, const int32_t *plain_slot, KswJob *plain_jobs,
                                 bwg_rescue_window *win);

}  // namespace bwg
