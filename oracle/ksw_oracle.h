/*
 * CPU oracle for the Smith-Waterman hot path (test infrastructure only).
 *
 * This is a plain-C restatement of the reference's scalar kernels, used by the test suite and by
 * the CPU-baseline leg of bench.py as the checker. Nothing in the product library links or calls
 * it.  Every function cites the reference file:line it follows (paths relative to the reference
 * checkout root).
 *
 *   orc_ksw_extend2        crates/bwa-extend/src/sw.rs:203-493   (ksw_extend2)
 *   orc_clamp_band         crates/bwa-extend/src/sw.rs:288-308,
 *                          crates/bwa-neon/src/batched.rs:1305-1329
 *   orc_tracker_*          crates/bwa-extend/src/sw.rs:789-852   (SuboptimalTracker)
 *   orc_ksw_local_fwd      crates/bwa-extend/src/sw.rs:921-1075  (ksw_local_fwd)
 *   orc_ksw_align2         crates/bwa-extend/src/sw.rs:1108-1185 (ksw_align2)
 *   orc_padded_qlen        crates/bwa-neon/src/matesw.rs:676-686 (ksw_padded_qlen),
 *                          lane rule crates/bwa-mem/src/pe.rs:420-424
 *   orc_batched_align2_job crates/bwa-neon/src/matesw.rs:299-497 (batched_ksw_align2, per job)
 *   orc_ungapped_*         crates/bwa-neon/src/batched.rs:493-521, 602-697, 756-764 (ksw_ungapped.c)
 *   orc_extend_side        crates/bwa-mem/src/lib.rs:332-373     (extend_side; ksw_side.c)
 *   orc_run_side           crates/bwa-mem/src/across.rs:1357-1460 (run_side; ksw_side.c)
 *   orc_ksw_global2        crates/bwa-extend/src/sw.rs:515-740 (ksw_global2, push_cigar; ksw_global.c)
 *   orc_matesw_*           crates/bwa-mem/src/pe.rs:192-269, 596-679, 1040-1113 and
 *                          crates/bwa-index/src/bntseq.rs:237-243, 337-367 (mate rescue windows;
 *                          ksw_matesw.c)
 *
 * A second, independent restatement in the lockstep (SoA, many jobs per pass) form lives in
 * ksw_lockstep.c, and brute-force checkers in ksw_brute.c.
 */
#ifndef KSW_ORACLE_H
#define KSW_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { int32_t score, qle, tle, gtle, gscore, max_off; } orc_ext_result;
typedef struct { int32_t score, qb, qe, tb, te, score2, te2; } orc_ksw_result;
typedef struct { int32_t score, te, qe, score2, te2; } orc_fwd_result;

/* Band clamp of ksw_extend2: f64 formula exactly as the reference evaluates it. */
int32_t orc_clamp_band(int32_t w, int32_t qlen, int32_t max_sc, int32_t end_bonus, int32_t o_ins,
                       int32_t e_ins, int32_t o_del, int32_t e_del);
/* Integer closed form of the same clamp (what the product library uses); compared exhaustively. */
int32_t orc_clamp_band_int(int32_t w, int32_t qlen, int32_t max_sc, int32_t end_bonus,
                           int32_t o_ins, int32_t e_ins, int32_t o_del, int32_t e_del);

/* ksw_extend2. `cells` (optional) accumulates the number of inner-loop cells evaluated. */
orc_ext_result orc_ksw_extend2(int32_t qlen, const uint8_t *query, int32_t tlen,
                               const uint8_t *target, int32_t m, const int8_t *mat, int32_t o_del,
                               int32_t e_del, int32_t o_ins, int32_t e_ins, int32_t w,
                               int32_t end_bonus, int32_t zdrop, int32_t h0, int64_t *cells);

/* Suboptimal tracker (score2/te2).  `buf` must hold at least (number of pushed rows) pairs. */
typedef struct { int32_t *score; int32_t *row; int32_t n; } orc_tracker;
void orc_tracker_push_row(orc_tracker *b, int32_t row, int32_t row_max, int32_t minsc);
void orc_tracker_finish(const orc_tracker *b, int32_t gmax, int32_t te, int32_t max_sc,
                        int32_t *score2, int32_t *te2);

/* ksw_local_fwd: query padded with score-0 columns to a multiple of `lanes`. */
orc_fwd_result orc_ksw_local_fwd(int32_t qlen, const uint8_t *query, int32_t tlen,
                                 const uint8_t *target, int32_t m, const int8_t *mat,
                                 int32_t o_del, int32_t e_del, int32_t o_ins, int32_t e_ins,
                                 int32_t minsc, int32_t endsc, int32_t max_sc, int32_t lanes,
                                 int64_t *cells);

int32_t orc_lanes_for(int32_t qlen, int32_t max_sc);
int32_t orc_padded_qlen(int32_t qlen, int32_t max_sc);

/* ksw_align2 (scalar, sw.rs semantics: both passes use the forward `lanes`, full target kept). */
orc_ksw_result orc_ksw_align2(int32_t qlen, const uint8_t *query, int32_t tlen,
                              const uint8_t *target, int32_t m, const int8_t *mat, int32_t o_del,
                              int32_t e_del, int32_t o_ins, int32_t e_ins, int32_t minsc,
                              int32_t max_sc, int32_t lanes, int64_t *cells);

/* batched_ksw_align2 for one job (matesw.rs semantics: each pass padded by its own length, the
 * reverse pass runs on the reversed prefix target[..=te] only). */
orc_ksw_result orc_batched_align2_job(int32_t qlen, const uint8_t *query, int32_t tlen,
                                      const uint8_t *target, int32_t m, const int8_t *mat,
                                      int32_t o_del, int32_t e_del, int32_t o_ins, int32_t e_ins,
                                      int32_t minsc, int32_t max_sc, int64_t *cells);

/* Batch drivers over a byte arena (the same job description the product C ABI takes), run on
 * `nthreads` host threads with a static partition.  Used by tests and by the CPU baseline. */
void orc_extend_batch(size_t n, const uint8_t *arena, const uint64_t *q_off, const uint32_t *q_len,
                      const uint64_t *t_off, const uint32_t *t_len, const int32_t *h0, int32_t m,
                      const int8_t *mat, int32_t o_del, int32_t e_del, int32_t o_ins,
                      int32_t e_ins, int32_t w, int32_t end_bonus, int32_t zdrop,
                      orc_ext_result *out, int64_t *cells_out, int nthreads);
/* orc_extend_batch with the evaluated-cell count of every job in job_cells[n]. */
void orc_extend_batch_cells(size_t n, const uint8_t *arena, const uint64_t *q_off, const uint32_t *q_len,
                            const uint64_t *t_off, const uint32_t *t_len, const int32_t *h0, int32_t m,
                            const int8_t *mat, int32_t o_del, int32_t e_del, int32_t o_ins, int32_t e_ins,
                            int32_t w, int32_t end_bonus, int32_t zdrop, orc_ext_result *out,
                            int32_t *job_cells, int nthreads);
void orc_align2_batch(size_t n, const uint8_t *arena, const uint64_t *q_off, const uint32_t *q_len,
                      const uint64_t *t_off, const uint32_t *t_len, int32_t m, const int8_t *mat,
                      int32_t o_del, int32_t e_del, int32_t o_ins, int32_t e_ins, int32_t minsc,
                      int32_t max_sc, int batched_semantics, orc_ksw_result *out,
                      int64_t *cells_out, int nthreads);

/* SIMD CPU baseline (ksw_simd.c): the reference's inter-sequence form on AVX2, 16 jobs per chunk,
 * the same results as orc_extend_batch and orc_align2_batch (batched semantics). */
void orc_simd_extend_batch(size_t n, const uint8_t *arena, const uint64_t *q_off, const uint32_t *q_len,
                           const uint64_t *t_off, const uint32_t *t_len, const int32_t *h0, int32_t m,
                           const int8_t *mat, int32_t o_del, int32_t e_del, int32_t o_ins, int32_t e_ins,
                           int32_t w, int32_t end_bonus, int32_t zdrop, orc_ext_result *out, int nthreads);
void orc_simd_align2_batch(size_t n, const uint8_t *arena, const uint64_t *q_off, const uint32_t *q_len,
                           const uint64_t *t_off, const uint32_t *t_len, int32_t m, const int8_t *mat,
                           int32_t o_del, int32_t e_del, int32_t o_ins, int32_t e_ins, int32_t minsc,
                           int32_t max_sc, orc_ksw_result *out, int nthreads);

/* Lockstep twins (ksw_lockstep.c): crates/bwa-neon/src/batched.rs:947-1251 and
 * crates/bwa-neon/src/matesw.rs:933-1232, `lanes` jobs per pass in SoA layout. */
void orc_lockstep_extend(size_t n, const uint8_t *const *queries, const int32_t *qlens,
                         const uint8_t *const *targets, const int32_t *tlens, const int32_t *h0,
                         int32_t m, const int8_t *mat, int32_t o_del, int32_t e_del,
                         int32_t o_ins, int32_t e_ins, int32_t w, int32_t end_bonus,
                         int32_t zdrop, int32_t lanes, orc_ext_result *out);
void orc_lockstep_local_fwd(size_t n, const uint8_t *const *queries, const int32_t *qlens,
                            const uint8_t *const *targets, const int32_t *tlens,
                            const int32_t *minsc, const int32_t *endsc, int32_t m,
                            const int8_t *mat, int32_t o_del, int32_t e_del, int32_t o_ins,
                            int32_t e_ins, int32_t max_sc, int32_t lanes, orc_fwd_result *out);

/* Brute-force checkers (ksw_brute.c). */
#define ORC_ENUM_NONE (-(1 << 30))
void orc_enum_prefix(int32_t qlen, const uint8_t *query, int32_t tlen, const uint8_t *target, const int8_t *mat,
                     int32_t o_del, int32_t e_del, int32_t o_ins, int32_t e_ins, int32_t w, int32_t origin,
                     int32_t *Mx, int32_t *Ex, int32_t *Fx);
/* Unbanded extension recurrence without band, z-drop or zero-row break: best score only
 * (crates/bwa-extend/src/sw.rs:1255-1305). */
int32_t orc_brute_extend_unbanded(int32_t qlen, const uint8_t *query, int32_t tlen,
                                  const uint8_t *target, int32_t m, const int8_t *mat,
                                  int32_t o_del, int32_t e_del, int32_t o_ins, int32_t e_ins,
                                  int32_t h0);
/* Textbook Gotoh local alignment (three matrices, gaps opened from H), best score only. */
int32_t orc_brute_gotoh_local(int32_t qlen, const uint8_t *query, int32_t tlen,
                              const uint8_t *target, int32_t m, const int8_t *mat, int32_t o_del,
                              int32_t e_del, int32_t o_ins, int32_t e_ins);
/* Exhaustive enumeration of local alignment paths on tiny inputs (qlen,tlen <= 7): best score,
 * the first target row reaching it, and the smallest query column reaching it in that row. */
void orc_brute_enum_local(int32_t qlen, const uint8_t *query, int32_t tlen, const uint8_t *target,
                          int32_t m, const int8_t *mat, int32_t o_del, int32_t e_del,
                          int32_t o_ins, int32_t e_ins, int32_t *score, int32_t *te, int32_t *qe);

/* Ungapped-diagonal fast path (ksw_ungapped.c). */
int orc_is_uniform_dna(const int8_t *mat, int32_t m);
int32_t orc_ungapped_xmax(const int8_t *mat, int32_t o_del, int32_t e_del, int32_t o_ins, int32_t e_ins);
int orc_ungapped_hit(int32_t qlen, const uint8_t *query, int32_t tlen, const uint8_t *target, int32_t h0,
                     int32_t a, int32_t b, int32_t xmax, int32_t zdrop, orc_ext_result *out);
void orc_ungapped_batch(size_t n, const uint8_t *arena, const uint64_t *q_off, const uint32_t *q_len,
                        const uint64_t *t_off, const uint32_t *t_len, const int32_t *h0, int32_t m,
                        const int8_t *mat, int32_t o_del, int32_t e_del, int32_t o_ins, int32_t e_ins,
                        int32_t zdrop, orc_ext_result *out, uint8_t *hit);

/* Band doubling around ksw_extend2 (ksw_side.c): one job's tries, and the batched rounds.  w_out
 * receives the accepted band. */
orc_ext_result orc_extend_side(int32_t qlen, const uint8_t *query, int32_t tlen, const uint8_t *target,
                               int32_t m, const int8_t *mat, int32_t o_del, int32_t e_del, int32_t o_ins,
                               int32_t e_ins, int32_t w, int32_t end_bonus, int32_t zdrop, int32_t h0,
                               int32_t prev0, int32_t max_band_try, int32_t *w_out);
void orc_run_side(size_t n, const uint8_t *arena, const uint64_t *q_off, const uint32_t *q_len,
                  const uint64_t *t_off, const uint32_t *t_len, const int32_t *h0, const int32_t *prev0,
                  int32_t m, const int8_t *mat, int32_t o_del, int32_t e_del, int32_t o_ins, int32_t e_ins,
                  int32_t w, int32_t end_bonus, int32_t zdrop, int32_t max_band_try, orc_ext_result *out,
                  int32_t *w_out, int nthreads);

/* Mate rescue windows (ksw_matesw.c).  Layouts match bwg_mate_anchor and bwg_rescue_window. */
typedef struct {
    int64_t rb;
    int32_t rid;
    uint32_t l_ms;
    uint64_t ms_off, ma_begin, ma_end;
} orc_mate_anchor;
typedef struct {
    int64_t rb, re;
    int32_t rid, flags;
} orc_rescue_window;
void orc_infer_dir(int64_t l_pac, int64_t b1, int64_t b2, int32_t *dir, int64_t *dist);
int64_t orc_depos(int64_t l_pac, int64_t pos, int *is_rev);
int32_t orc_pos2rid(int64_t l_pac, const int64_t *off, int32_t n, int64_t pos_f);
void orc_fetch_window(int64_t l_pac, const int64_t *off, const int32_t *len, int32_t n, int64_t rb, int64_t mid,
                      int64_t re, int64_t *rb_out, int64_t *re_out, int32_t *rid_out);
uint8_t orc_ref_base(const uint8_t *fwd, int64_t l_pac, int64_t p);
void orc_matesw_collect(int64_t l_pac, const int64_t *off, const int32_t *len, int32_t n_contigs,
                        const int32_t *low, const int32_t *high, const int32_t *failed, int32_t min_seed_len,
                        int64_t a_rb, int32_t a_rid, int64_t l_ms, const int64_t *ma_rb, size_t n_ma,
                        orc_rescue_window *win);
void orc_matesw_batch(int64_t l_pac, const uint8_t *fwd, const int64_t *off, const int32_t *len, int32_t n_contigs,
                      const int32_t *low, const int32_t *high, const int32_t *failed, int32_t min_seed_len,
                      size_t n, const orc_mate_anchor *anchors, const int64_t *ma_rb, const uint8_t *arena,
                      const int8_t *mat, int32_t o_del, int32_t e_del, int32_t o_ins, int32_t e_ins, int32_t minsc,
                      int32_t max_sc, orc_rescue_window *win, orc_ksw_result *out, int nthreads);

/* Banded global alignment with traceback (ksw_global.c).  cigar needs qlen + tlen + 1 words. */
int32_t orc_ksw_global2(int32_t qlen, const uint8_t *query, int32_t tlen, const uint8_t *target, int32_t m,
                        const int8_t *mat, int32_t o_del, int32_t e_del, int32_t o_ins, int32_t e_ins, int32_t w,
                        uint32_t *cigar, int32_t *n_cigar);
void orc_global_batch(size_t n, const uint8_t *arena, const uint64_t *q_off, const uint32_t *q_len,
                      const uint64_t *t_off, const uint32_t *t_len, const int32_t *w, const int8_t *mat,
                      int32_t o_del, int32_t e_del, int32_t o_ins, int32_t e_ins, int32_t *score,
                      const uint64_t *cig_off, uint32_t *cigar, int32_t *n_cigar, int nthreads);
int32_t orc_brute_global(int32_t qlen, const uint8_t *query, int32_t tlen, const uint8_t *target, const int8_t *mat,
                         int32_t o_del, int32_t e_del, int32_t o_ins, int32_t e_ins);

#ifdef __cplusplus
}
#endif
#endif
