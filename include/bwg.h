/*
 * bwg: MI355X (gfx950) Smith-Waterman backend for bwa-mem4's extension and mate-rescue hot path.
 *
 * C ABI of libbwg.so.  Plain pointers and sizes only.  Every entry point returns 0 (BWG_OK) on
 * success and a negative errno-style code otherwise; nothing falls back to a CPU path.
 *
 * The two compute entry points replace, one for one, the reference's plug points for this path:
 *
 *   bwg_extend_batch      SwBackend::extend_batch   crates/bwa-extend/src/lib.rs:154-175
 *                         (called at crates/bwa-mem/src/across.rs:1402-1413; the production
 *                         implementor is chosen at crates/bwa-cli/src/cmd_mem.rs:2214)
 *   bwg_extend            SwBackend::extend         crates/bwa-extend/src/lib.rs:115-129
 *   bwg_ksw_align2_batch  batched_ksw_align2        crates/bwa-neon/src/matesw.rs:299-309
 *                         (called at crates/bwa-mem/src/pe.rs:1090-1100)
 *   bwg_name              SwBackend::name           crates/bwa-extend/src/lib.rs:85
 *
 * Jobs are described as offsets/lengths into one caller-owned byte arena holding 2-bit base codes
 * (0=A 1=C 2=G 3=T 4=N), which is how a Rust caller can pass its borrowed slices without copying.
 * Results are written into caller-allocated arrays, one per job, in job order.
 */
#ifndef BWG_H
#define BWG_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    BWG_OK = 0,
    BWG_EIO = -5,      /* a HIP runtime call or kernel failed */
    BWG_E2BIG = -7,    /* a job exceeds the size limits stated below, or an output buffer is too small */
    BWG_EAGAIN = -11,  /* serving: the HBM budget cannot admit the chunk right now */
    BWG_ENOMEM = -12,  /* device or host allocation failed */
    BWG_ENODEV = -19,  /* no usable gfx950 device */
    BWG_EINVAL = -22   /* a precondition of the reference kernel is violated */
};

/* ExtendResult, crates/bwa-extend/src/sw.rs:138-162 (field order kept). */
typedef struct { int32_t score, qle, tle, gtle, gscore, max_off; } bwg_ext_result;
/* KswAlignResult, crates/bwa-extend/src/sw.rs:749-774 (field order kept). */
typedef struct { int32_t score, qb, qe, tb, te, score2, te2; } bwg_ksw_result;

/* Shared scoring of one extend_batch call: the arguments of SwBackend::extend_batch other than
 * the jobs.  m must be 5; mat is row-major mat[t*5+q]. */
typedef struct {
    int32_t m;
    int8_t mat[25];
    int32_t o_del, e_del, o_ins, e_ins, w, end_bonus, zdrop;
} bwg_ext_params;

/* Shared scoring of one batched_ksw_align2 call. */
typedef struct {
    int32_t m;
    int8_t mat[25];
    int32_t o_del, e_del, o_ins, e_ins, minsc, max_sc;
} bwg_ksw_params;

typedef struct bwg_ctx bwg_ctx;

/* Create a context on HIP device `device` (its own stream and device buffers).  Returns
 * BWG_ENODEV when the device is absent or is not gfx950. */
int bwg_create(int device, bwg_ctx **out);
void bwg_destroy(bwg_ctx *ctx);
/* "hip-gfx950" (SwBackend::name). */
const char *bwg_name(const bwg_ctx *ctx);
const char *bwg_strerror(int code);

/* Validation applied to every job (BWG_EINVAL on failure), the preconditions of ksw_extend2 and
 * ksw_align2: m == 5; e_del, e_ins >= 1; base codes <= 4; extension h0 > 0 and qlen >= 1;
 * rescue max_sc >= 1.  Size limits (BWG_E2BIG): query and target up to 2^24 bases each, with the
 * largest score (extension h0 + qlen*max_sc + end_bonus, rescue (qlen+16)*max(mat)) and the largest
 * gap cost (max(e_del,e_ins) * (max(qlen,tlen)+16) + max(o_del,o_ins)) below 2^30, so that no
 * int32 quantity of the recurrences can overflow.  Jobs beyond the binned kernels' shapes run in
 * the plain kernels (per-job state in global memory), with identical results.
 *
 * Threading.  The synchronous entry points (bwg_extend_batch, bwg_extend, bwg_ksw_align2_batch,
 * bwg_extend_side_batch, bwg_global_batch, bwg_matesw_batch) may be called concurrently from any
 * number of host threads on one context: each call runs on one of the context's execution lanes
 * (BWG_OPT_LANES, default 4; each its own set of HIP streams), so up to that many calls stage,
 * run and copy back at the same time and the rest wait for a lane.  Extension and rescue calls
 * of up to 65536 jobs that arrive while another synchronous call is running are instead merged
 * with the other calls waiting at that moment into one device batch (BWG_OPT_COALESCE, default
 * on; the context's own server, see bwg_server_*): one batch of many small calls costs the device
 * about what one of them does, and every job's result is the same as in a call of its own.  The staged-batch calls
 * (bwg_batch_*) share the context's own streams and run one at a time.  bwg_set_option may be
 * called at any time; a batch plans with the options in force when it was staged (staged batches)
 * or called (synchronous entry points). */

/* Banded seed extension (ksw_extend2 semantics) for n jobs sharing `p`.  Thread-safe; may be
 * called concurrently from many host threads. */
int bwg_extend_batch(bwg_ctx *ctx, const bwg_ext_params *p, size_t n, const uint8_t *arena,
                     size_t arena_len, const uint64_t *q_off, const uint32_t *q_len,
                     const uint64_t *t_off, const uint32_t *t_len, const int32_t *h0,
                     bwg_ext_result *out);

/* bwg_extend_batch without any device allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, semantics and results as bwg_extend_batch;
 * the device memory beyond the caller's inputs and outputs (the staged arena, descriptors and
 * results, the bin queues, the plain kernel's state) comes from one 64 MiB array per GPU that the
 * library holds for the life of the process.  A call needing more (about 250k long-tail jobs)
 * returns BWG_E2BIG.  It runs the batched form on the context's own streams, plans with the
 * autotuner's table as it stands (it never tunes) and replays no graph.  Calls on one GPU share
 * the array, which keeps the last call's contents: the library serialises them (overlapping calls
 * would race on it). */
int bwg_extend_batch_noalloc(bwg_ctx *ctx, const bwg_ext_params *p, size_t n, const uint8_t *arena,
                             size_t arena_len, const uint64_t *q_off, const uint32_t *q_len,
                             const uint64_t *t_off, const uint32_t *t_len, const int32_t *h0,
                             bwg_ext_result *out);

/* One extension (SwBackend::extend). */
int bwg_extend(bwg_ctx *ctx, const bwg_ext_params *p, const uint8_t *query, uint32_t qlen,
               const uint8_t *target, uint32_t tlen, int32_t h0, bwg_ext_result *out);

/* Mate-rescue local alignment (batched_ksw_align2 semantics) for n jobs sharing `p`. */
int bwg_ksw_align2_batch(bwg_ctx *ctx, const bwg_ksw_params *p, size_t n, const uint8_t *arena,
                         size_t arena_len, const uint64_t *q_off, const uint32_t *q_len,
                         const uint64_t *t_off, const uint32_t *t_len, bwg_ksw_result *out);

/* ---- execution form, tuning and statistics --------------------------------------------------
 * BWG_OPT_FORM selects how the kernels are mapped: BWG_FORM_PLAIN is one thread per job in
 * arrival order with per-job state in global memory and one launch per batch; BWG_FORM_BATCHED
 * (default) bins jobs by length and score range and runs each bin with the mapping the autotuner
 * picked.  Both give identical results.
 *
 * Autotuning (BWG_OPT_AUTOTUNE, default 1): the first batch with at least 4096 jobs of a length
 * class times the class's candidate mappings on a sample of those jobs (one lane per job or 16
 * lanes per job; the group kernel's columns per lane; 16- or 32-bit stored DP state) and the
 * context keeps the fastest for that scoring.  Setting the option clears the table.  With
 * autotuning off, the mapping follows BWG_OPT_EXT_THREAD_MAX_Q, the K_STEP options and
 * BWG_OPT_EXT_SCORE_WIDTH / BWG_OPT_KSW_SCORE_WIDTH, which force each candidate for tests. */
enum { BWG_FORM_PLAIN = 0, BWG_FORM_BATCHED = 1 };
enum {
    BWG_OPT_FORM = 1,
    BWG_OPT_CONCURRENT_BINS = 2,  /* batched: run a batch's independent bins concurrently (default 1) */
    BWG_OPT_GRAPHS = 3,           /* batched: replay repeated runs of a batch from a hipGraph (default 1) */
    BWG_OPT_EXT_THREAD_MAX_Q = 4, /* batched: longest query run one job per lane (0..32, default 32) */
    BWG_OPT_AUTOTUNE = 5,         /* batched: time candidate mappings per length class (default 1) */
    BWG_OPT_EXT_K_STEP = 6,       /* untuned extension group bins use one K above the minimal (default 0) */
    BWG_OPT_KSW_K_STEP = 7,       /* untuned rescue bins use one K above the minimal (default 0) */
    BWG_OPT_EXT_SCORE_WIDTH = 8,  /* extension DP state bits: 16 where exact (default), or 32 always */
    BWG_OPT_UNGAPPED = 9,         /* batched extension: ungapped-diagonal fast path first (default 1) */
    BWG_OPT_KSW_SCORE_WIDTH = 10, /* rescue DP state bits: 16 where exact (default), or 32 always */
    BWG_OPT_LANES = 11,           /* execution lanes of the synchronous entry points (1..16, default 4) */
    BWG_OPT_POOL_CACHE_MB = 12,   /* device blocks the context keeps cached for reuse, MiB (default 8192) */
    BWG_OPT_POOL_PEAK_RESET = 13, /* any value: restart BWG_STAT_POOL_PEAK at the bytes now in use */
    BWG_OPT_ITEM_TIMING = 14,     /* batched: time every launch with HIP events (replays no graph; default 0) */
    BWG_OPT_COALESCE = 15,        /* merge concurrent small synchronous calls into one device batch (default 1) */
    BWG_OPT_HOST_THREADS = 16,    /* host threads for a call's validation and planning (0: auto, calls of 64k+ jobs; 1..16: that many, 1024+ jobs a thread) */
    BWG_OPT_EXT_GROUP_LANES = 17, /* untuned extension group bins: 16 (default) or 8 lanes per job, eight jobs a wavefront */
    BWG_OPT_KSW_GROUP_LANES = 18, /* untuned packed rescue bins: 16 (default) or 8 lanes per job, eight jobs a wavefront */
    BWG_OPT_BIN_STREAMS = 19,     /* batched: streams a stage's concurrent bins are dealt onto (1..4, default 4) */
    BWG_STAT_LAUNCHES = 100,    /* device kernel launches issued by this context (proof the GPU ran) */
    BWG_STAT_JOBS = 101,        /* jobs completed on the device */
    BWG_STAT_LAST_KERNEL_NS = 102, /* device time of the last batch's kernels (HIP events) */
    BWG_STAT_LAST_MAIN_NS = 103,   /* device time of the last batch's dominant DP kernel */
    BWG_STAT_LAST_MAIN_LAUNCHES = 104, /* launches of that dominant kernel in the last batch */
    BWG_STAT_POOL_IN_USE = 105,    /* device bytes the context's batches hold now */
    BWG_STAT_POOL_CACHED = 106,    /* device bytes freed by batches and cached for reuse */
    BWG_STAT_POOL_PEAK = 107,      /* largest BWG_STAT_POOL_IN_USE since creation or the last reset */
    /* host wall time of the phases of the last synchronous extension or rescue call, ns: host-side
     * validation; descriptor build and host-to-device copies (a large arena uploads from a helper
     * thread while the call plans: only the part of that upload planning did not hide counts
     * here); planning (binning, autotuning, plan uploads); enqueueing the kernels and waiting for
     * them; result copy to the host; the whole call.  Calls of 64k jobs or more split validation,
     * descriptor build and binning over host threads (BWG_OPT_HOST_THREADS). */
    BWG_STAT_PHASE_VALIDATE_NS = 110,
    BWG_STAT_PHASE_STAGE_NS = 111,
    BWG_STAT_PHASE_PLAN_NS = 112,
    BWG_STAT_PHASE_KERNELS_NS = 113,
    BWG_STAT_PHASE_FETCH_NS = 114,
    BWG_STAT_PHASE_TOTAL_NS = 115
};
int bwg_set_option(bwg_ctx *ctx, int key, int64_t value);
int64_t bwg_get_stat(const bwg_ctx *ctx, int key);
/* The autotuner's table as text, one line per tuned class with every candidate's time per job and
 * the chosen one starred.  Writes at most len-1 bytes plus a terminating 0; returns the full
 * length. */
int bwg_tune_report(const bwg_ctx *ctx, char *buf, size_t len);

/* Diagnostics, no counterpart in the reference: copy `bytes` (a multiple of 16) between two fresh
 * device buffers `reps` times with a streaming kernel and report the copy rate, counting each byte
 * read and each byte written.  Its known traffic calibrates hardware memory-traffic counters. */
int bwg_hbm_copy(bwg_ctx *ctx, size_t bytes, int reps, double *gb_per_s);

/* ---- device-resident batches ----------------------------------------------------------------
 * Stage a batch into device memory once (host-to-device copy), then run it any number of times
 * with its inputs resident in HBM, and copy results back.  bwg_extend_batch and
 * bwg_ksw_align2_batch are exactly stage + run + results + destroy. */
typedef struct bwg_batch bwg_batch;
int bwg_batch_stage_ext(bwg_ctx *ctx, const bwg_ext_params *p, size_t n, const uint8_t *arena,
                        size_t arena_len, const uint64_t *q_off, const uint32_t *q_len,
                        const uint64_t *t_off, const uint32_t *t_len, const int32_t *h0,
                        bwg_batch **out);
int bwg_batch_stage_ksw(bwg_ctx *ctx, const bwg_ksw_params *p, size_t n, const uint8_t *arena,
                        size_t arena_len, const uint64_t *q_off, const uint32_t *q_len,
                        const uint64_t *t_off, const uint32_t *t_len, bwg_batch **out);
/* Enqueue the batch's kernels on the context stream and wait for them. */
int bwg_batch_run(bwg_ctx *ctx, bwg_batch *b);
/* Copy results (bwg_ext_result or bwg_ksw_result, n of them) to host memory. */
int bwg_batch_results(bwg_ctx *ctx, bwg_batch *b, void *out);
void bwg_batch_destroy(bwg_ctx *ctx, bwg_batch *b);
/* Diagnostics, no counterpart in the reference.  bwg_batch_report: the launches of the batch's
 * last batched run, one text line each ("<index> round=<band-retry round> stage=<s> <kernel> jobs=<n>
 * grid=<workgroups> us=<device time, -1 unless BWG_OPT_ITEM_TIMING was on>"); writes at most len-1
 * bytes plus a terminating 0 and returns the full length.  bwg_batch_job_bins: for each job, the
 * index of the launch whose bin it was planned into (extension DP bins, including jobs the
 * ungapped pre-pass then resolved; rescue forward bins; -1 for none). */
int bwg_batch_report(bwg_ctx *ctx, bwg_batch *b, char *buf, size_t len);
int bwg_batch_job_bins(bwg_ctx *ctx, bwg_batch *b, int32_t *bin);
/* Extension jobs of the batch's last batched run that the ungapped-diagonal fast path resolved
 * without the banded DP (the reference's batched extension does the same, crates/bwa-neon/src/
 * batched.rs:716-798); 0 for rescue batches, the plain form, or with BWG_OPT_UNGAPPED off. */
int bwg_batch_ungapped_hits(bwg_ctx *ctx, bwg_batch *b, uint64_t *hits);

/* ---- one-sided extension with band doubling ------------------------------------------------------
 * Replaces the reference's batched band-retry driver run_side (crates/bwa-mem/src/across.rs:1357-1460),
 * whose per-read twin is extend_side (crates/bwa-mem/src/lib.rs:332-373): round r extends the jobs
 * still pending at band p->w << r (p->end_bonus is the side's clipping penalty), and a job is
 * accepted when its score equals its previous score, when its best cell's diagonal offset max_off is
 * below (w >> 1) + (w >> 2), or in the last of max_band_try rounds (bwa: 2).  prev[k] is the
 * job's score before the first round (-1 for a left extension; the region's score for a right one,
 * as the reference's across.rs:548-556 describes).  All rounds run on the device in one launch
 * sequence: the acceptance test between rounds compacts the rejected jobs into the next round's
 * bins, with no host round trip.  out[k] is the accepted round's result and w_out[k] its band, the
 * value the caller folds into the region's w.  The region bookkeeping (local versus to-end choice,
 * across.rs:1446-1512) stays with the caller.  Side batches always run the batched engine. */
int bwg_extend_side_batch(bwg_ctx *ctx, const bwg_ext_params *p, int max_band_try, size_t n,
                          const uint8_t *arena, size_t arena_len, const uint64_t *q_off,
                          const uint32_t *q_len, const uint64_t *t_off, const uint32_t *t_len,
                          const int32_t *h0, const int32_t *prev, bwg_ext_result *out, int32_t *w_out);
/* Staged form of bwg_extend_side_batch: run with bwg_batch_run, results with bwg_batch_results,
 * accepted bands with bwg_batch_side_bands. */
int bwg_batch_stage_side(bwg_ctx *ctx, const bwg_ext_params *p, int max_band_try, size_t n,
                         const uint8_t *arena, size_t arena_len, const uint64_t *q_off,
                         const uint32_t *q_len, const uint64_t *t_off, const uint32_t *t_len,
                         const int32_t *h0, const int32_t *prev, bwg_batch **out);
int bwg_batch_side_bands(bwg_ctx *ctx, bwg_batch *b, int32_t *w_out);

/* ---- banded global alignment with traceback -------------------------------------------------
 * Replaces ksw_global2 (crates/bwa-extend/src/sw.rs:561-740), called per alignment region at
 * crates/bwa-mem/src/cigar.rs:446 to turn a settled region into its CIGAR: end-to-end alignment of
 * all of the query against all of the target inside the band |j - i| <= w, gaps opened from
 * aligned pairs, unreachable cells at MINUS_INF, ties M over E over F, traceback from the band's
 * last cell.  Each job has its own band w[k] (the caller's clamp, cigar.rs:417-444), which must
 * cover the length difference (w[k] >= |qlen - tlen|, as that clamp guarantees).  Queries up to
 * 512 bases run in the binned group kernels; longer ones one job per thread, with a code byte per
 * band cell in device memory (BWG_E2BIG when tlen * min(qlen, 2w + 1) exceeds 2^30, or a side
 * exceeds 2^24 bases).
 * score[k] is the reference's score; the CIGAR words are len << 4 | op (0 = M, 1 = I, 2 = D), the
 * runs of job k following those of job k-1 in `cigar`, n_cigar[k] of them. */
typedef struct { int32_t m; int8_t mat[25]; int32_t o_del, e_del, o_ins, e_ins; } bwg_global_params;
/* BWG_E2BIG when cigar_cap is smaller than the total number of runs (n_cigar is filled anyway). */
int bwg_global_batch(bwg_ctx *ctx, const bwg_global_params *p, size_t n, const uint8_t *arena,
                     size_t arena_len, const uint64_t *q_off, const uint32_t *q_len,
                     const uint64_t *t_off, const uint32_t *t_len, const int32_t *w, int32_t *score,
                     uint32_t *n_cigar, uint32_t *cigar, size_t cigar_cap);
/* Staged form: run with bwg_batch_run, scores (n int32) with bwg_batch_results, then the run
 * counts and the runs. */
int bwg_batch_stage_global(bwg_ctx *ctx, const bwg_global_params *p, size_t n, const uint8_t *arena,
                           size_t arena_len, const uint64_t *q_off, const uint32_t *q_len,
                           const uint64_t *t_off, const uint32_t *t_len, const int32_t *w,
                           bwg_batch **out);
int bwg_batch_cigar_lengths(bwg_ctx *ctx, bwg_batch *b, uint32_t *n_cigar);
int bwg_batch_cigars(bwg_ctx *ctx, bwg_batch *b, uint32_t *cigar, size_t cigar_cap);

/* ---- mate rescue with device-resident windows -----------------------------------------------
 * Replaces one round of the reference's batched mate rescue (crates/bwa-mem/src/pe.rs:1040-1113):
 * matesw_collect (pe.rs:596-679) with its reference fetch bns_fetch_seq (pe.rs:211-269), the flat
 * batched_ksw_align2 call (matesw.rs:299-309), and nothing of matesw_apply (pe.rs:716), which stays
 * with the caller.  The reference (bwg_set_reference) lives on the device as both strands, one byte
 * per base (2 * l_pac bytes); windows and their sequences never leave the device.
 *
 * Per anchor (one region `a` of one read, with the mate's sequence) and orientation r = 0..3 the
 * device computes what matesw_collect does: the skip test (pes[r].failed, or a mate region already
 * at a proper distance in direction r, mem_infer_dir pe.rs:192-209), the insert window from
 * pes[r].low/high and the mate length, the clamp to [0, 2*l_pac), the contig clamp around the
 * window's midpoint (mirrored on the reverse strand), and the rid and min_seed_len test; then it
 * aligns the mate (reverse-complemented when the orientation says so) against the window with
 * batched_ksw_align2 semantics. */
typedef struct { int64_t offset; int32_t len; } bwg_contig;          /* forward-strand annotation */
typedef struct { int32_t low, high, failed; } bwg_pe_orient;        /* PeStat, pe.rs:143-158 */
typedef struct {
    int64_t rb;          /* the anchor region's start in the doubled 2*l_pac space */
    int32_t rid;         /* its contig */
    uint32_t l_ms;       /* mate length */
    uint64_t ms_off;     /* mate sequence (codes 0..4, as read) in the arena */
    uint64_t ma_begin;   /* the mate's existing regions: ma_rb[ma_begin .. ma_end) */
    uint64_t ma_end;
} bwg_mate_anchor;
enum { BWG_WIN_VALID = 1, BWG_WIN_REV = 2, BWG_WIN_SKIPPED = 4 };
typedef struct {
    int64_t rb, re;      /* the clamped window in 2*l_pac space (valid windows) */
    int32_t rid;         /* contig of the window's midpoint, -1 outside every contig */
    int32_t flags;       /* BWG_WIN_*: a job ran / the mate was reverse-complemented / r skipped */
} bwg_rescue_window;
/* Upload the forward strand (codes 0..4, l_pac bytes) and its contigs (ascending offsets, the
 * reference's bns anns); the device builds the reverse-complement half.  Replaces any previous
 * reference of the context. */
int bwg_set_reference(bwg_ctx *ctx, const uint8_t *fwd, int64_t l_pac, const bwg_contig *contigs,
                      int32_t n_contigs);
/* One rescue round for n anchors: 4n windows and 4n results, index 4*k + r.  Jobs whose
 * window is not valid get the default result {0,-1,-1,-1,-1,-1,-1}.  p->minsc and p->max_sc are the
 * caller's min_seed_len * a and a (pe.rs:1096-1098); min_seed_len is the window-length floor. */
int bwg_matesw_batch(bwg_ctx *ctx, const bwg_ksw_params *p, int32_t min_seed_len,
                     const bwg_pe_orient pes[4], size_t n, const bwg_mate_anchor *anchors,
                     const int64_t *ma_rb, size_t n_ma, const uint8_t *arena, size_t arena_len,
                     bwg_rescue_window *win, bwg_ksw_result *out);
/* Staged form: run with bwg_batch_run, results (4n) with bwg_batch_results, windows with
 * bwg_batch_windows. */
int bwg_batch_stage_matesw(bwg_ctx *ctx, const bwg_ksw_params *p, int32_t min_seed_len,
                           const bwg_pe_orient pes[4], size_t n, const bwg_mate_anchor *anchors,
                           const int64_t *ma_rb, size_t n_ma, const uint8_t *arena, size_t arena_len,
                           bwg_batch **out);
int bwg_batch_windows(bwg_ctx *ctx, bwg_batch *b, bwg_rescue_window *win);

/* ---- serving ----------------------------------------------------------------------------------
 * Continuous batching of alignment work from concurrent producers.  No counterpart in the
 * reference: its callers run extension and rescue per rayon chunk
 * (crates/bwa-cli/src/cmd_mem.rs:2209-2216, crates/bwa-mem/src/pe.rs:1090-1100); a server lets those
 * callers submit their chunks from every worker thread and have them merged into device-sized
 * epochs.  A chunk holds jobs of one kind sharing one scoring.  Its inputs must stay valid and
 * unmodified, and its `out` array must stay allocated, until bwg_server_wait / bwg_server_poll
 * report completion.  Targets index the chunk arena, or, with BWG_TARGET_FROM_REF set in t_off,
 * the server's device-resident reference (bwg_server_set_reference); BWG_TARGET_REVERSE then
 * reads that reference window back to front (left extensions). */
#define BWG_TARGET_FROM_REF (1ull << 62)
#define BWG_TARGET_REVERSE (1ull << 63)
enum { BWG_CHUNK_EXTEND = 0, BWG_CHUNK_RESCUE = 1 };
typedef struct {
    int32_t kind;                /* BWG_CHUNK_EXTEND or BWG_CHUNK_RESCUE */
    const bwg_ext_params *ext;   /* scoring, kind == BWG_CHUNK_EXTEND */
    const bwg_ksw_params *ksw;   /* scoring, kind == BWG_CHUNK_RESCUE */
    size_t n;
    const uint8_t *arena;
    size_t arena_len;
    const uint64_t *q_off;
    const uint32_t *q_len;
    const uint64_t *t_off;
    const uint32_t *t_len;
    const int32_t *h0;           /* extension only */
    void *out;                   /* n bwg_ext_result or bwg_ksw_result */
} bwg_chunk;
typedef struct {
    uint64_t chunks, jobs, epochs;
    uint64_t peak_device_bytes;  /* largest admitted estimate of the epochs in flight (<= budget) */
    uint64_t peak_pool_bytes;    /* largest device bytes the context's batches actually held */
    double seconds, p50_ms, p99_ms, mean_ms, max_ms;  /* chunk latency: submit to results available */
} bwg_server_counters;
typedef struct bwg_server bwg_server;
/* hbm_budget_bytes bounds the device memory of the epochs in flight: a chunk is admitted when an
 * upper bound of what its batch allocates (staging, plan, plain-kernel state, row maxima, autotuner
 * samples) fits beside the epochs already running; n_slots (1..8) epochs can be on the device at
 * once. */
int bwg_server_create(bwg_ctx *ctx, uint64_t hbm_budget_bytes, int n_slots, bwg_server **out);
/* Upload the reference (codes 0..4) that BWG_TARGET_FROM_REF targets index; only while idle. */
int bwg_server_set_reference(bwg_server *s, const uint8_t *ref, uint64_t len);
/* Validate and queue a chunk (any thread).  BWG_E2BIG if it alone exceeds the budget. */
int bwg_server_submit(bwg_server *s, const bwg_chunk *chunk, uint64_t *ticket);
/* BWG_EAGAIN while the chunk is pending, else its final status; the ticket is then released,
 * unless a bwg_server_wait is in progress on it (that wait releases it). */
int bwg_server_poll(bwg_server *s, uint64_t ticket);
/* Block until the chunk completes; returns its status and releases the ticket.  One wait per
 * ticket: a second concurrent wait returns BWG_EINVAL. */
int bwg_server_wait(bwg_server *s, uint64_t ticket);
/* Counters and chunk-latency percentiles since creation or the last reset. */
int bwg_server_stats(bwg_server *s, bwg_server_counters *out, int reset);
/* Drain every queued chunk, then stop. */
void bwg_server_destroy(bwg_server *s);

#ifdef __cplusplus
}
#endif
#endif /* BWG_H */
