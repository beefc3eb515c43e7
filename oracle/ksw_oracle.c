/*
 * CPU oracle: plain-C restatement of the reference's scalar Smith-Waterman kernels.
 *
 * TEST INFRASTRUCTURE ONLY.  The product library never links, loads or calls this file; only the
 * test suite, the smoke check and bench.py's CPU-baseline leg do, as the checker.
 *
 * All arithmetic is int32 exactly as in the reference; the only floating point is the band clamp,
 * which the reference evaluates in f64 and which is reproduced here in f64 (orc_clamp_band) next
 * to its integer closed form (orc_clamp_band_int).
 */
#include "ksw_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* crates/bwa-extend/src/sw.rs:288-308 and crates/bwa-neon/src/batched.rs:1305-1329: the widest
 * gap that can pay for itself, computed as (int)((double)(qlen*max_sc + end_bonus - o)/e + 1.0),
 * floored at 1, for insertions and then deletions. */
int32_t orc_clamp_band(int32_t w, int32_t qlen, int32_t max_sc, int32_t end_bonus, int32_t o_ins,
                       int32_t e_ins, int32_t o_del, int32_t e_del) {
    double num_i = (double)qlen * (double)max_sc + (double)end_bonus - (double)o_ins;
    int32_t max_ins = (int32_t)(num_i / (double)e_ins + 1.0);
    if (max_ins < 1) max_ins = 1;
    if (w > max_ins) w = max_ins;
    double num_d = (double)qlen * (double)max_sc + (double)end_bonus - (double)o_del;
    int32_t max_del = (int32_t)(num_d / (double)e_del + 1.0);
    if (max_del < 1) max_del = 1;
    if (w > max_del) w = max_del;
    return w;
}

/* Integer form: with n = qlen*max_sc + end_bonus - o and e >= 1, trunc(n/e + 1.0) floored at 1
 * equals (n <= 0 ? 1 : n/e + 1). */
static int32_t clamp_term(int64_t n, int32_t e) { return n <= 0 ? 1 : (int32_t)(n / e + 1); }
int32_t orc_clamp_band_int(int32_t w, int32_t qlen, int32_t max_sc, int32_t end_bonus,
                           int32_t o_ins, int32_t e_ins, int32_t o_del, int32_t e_del) {
    int64_t base = (int64_t)qlen * max_sc + end_bonus;
    int32_t a = clamp_term(base - o_ins, e_ins), b = clamp_term(base - o_del, e_del);
    if (w > a) w = a;
    if (w > b) w = b;
    return w;
}

static int32_t mat_max(int32_t m, const int8_t *mat) {
    int32_t mx = mat[0];
    for (int32_t k = 1; k < m * m; ++k)
        if (mat[k] > mx) mx = mat[k];
    return mx;
}

/* crates/bwa-extend/src/sw.rs:203-493 (ksw_extend2). */
orc_ext_result orc_ksw_extend2(int32_t qlen, const uint8_t *query, int32_t tlen,
                               const uint8_t *target, int32_t m, const int8_t *mat, int32_t o_del,
                               int32_t e_del, int32_t o_ins, int32_t e_ins, int32_t w,
                               int32_t end_bonus, int32_t zdrop, int32_t h0, int64_t *cells) {
    const int32_t oe_del = o_del + e_del, oe_ins = o_ins + e_ins;
    /* query profile, sw.rs:233-244: prof[c*qlen + j] = mat[c*m + query[j]] */
    int8_t *prof = (int8_t *)malloc((size_t)qlen * m + 1);
    for (int32_t c = 0; c < m; ++c)
        for (int32_t j = 0; j < qlen; ++j) prof[c * qlen + j] = mat[c * m + query[j]];
    /* eh arrays over absolute query columns 0..qlen, zero-initialised (sw.rs:263-264) */
    int32_t *H = (int32_t *)calloc((size_t)qlen + 2, sizeof(int32_t));
    int32_t *E = (int32_t *)calloc((size_t)qlen + 2, sizeof(int32_t));
    /* row -1: insertion ramp (sw.rs:269-279) */
    H[0] = h0;
    H[1] = h0 > oe_ins ? h0 - oe_ins : 0;
    for (int32_t j = 2; j <= qlen && H[j - 1] > e_ins; ++j) H[j] = H[j - 1] - e_ins;
    /* band clamp (sw.rs:288-308) */
    w = orc_clamp_band(w, qlen, mat_max(m, mat), end_bonus, o_ins, e_ins, o_del, e_del);

    int32_t best = h0, best_i = -1, best_j = -1, gbest_i = -1, gscore = -1, max_off = 0;
    int32_t beg = 0, end = qlen;
    int64_t ncell = 0;
    for (int32_t i = 0; i < tlen; ++i) {
        int32_t f = 0, rmax = 0, rj = -1;
        const int8_t *q = prof + (size_t)target[i] * qlen;
        /* row prologue (sw.rs:342-360) */
        if (beg < i - w) beg = i - w;
        if (end > i + w + 1) end = i + w + 1;
        if (end > qlen) end = qlen;
        int32_t h1 = 0;
        if (beg == 0) {
            h1 = h0 - (o_del + e_del * (i + 1));
            if (h1 < 0) h1 = 0;
        }
        /* cells (sw.rs:363-414) */
        int32_t j = beg;
        for (; j < end; ++j) {
            int32_t diag = H[j], e = E[j];
            H[j] = h1;
            int32_t mm = diag != 0 ? diag + q[j] : 0;
            int32_t h = mm > e ? mm : e;
            if (f > h) h = f;
            h1 = h;
            if (!(rmax > h)) { rj = j; rmax = h; } /* ties go to the later column */
            int32_t t = mm - oe_del;
            if (t < 0) t = 0;
            e -= e_del;
            E[j] = e > t ? e : t;
            t = mm - oe_ins;
            if (t < 0) t = 0;
            f -= e_ins;
            f = f > t ? f : t;
        }
        ncell += end > beg ? end - beg : 0;
        /* row epilogue (sw.rs:418-480) */
        H[end] = h1;
        E[end] = 0;
        if (j == qlen && gscore <= h1) { gbest_i = i; gscore = h1; }
        if (rmax == 0) break;
        if (rmax > best) {
            best = rmax;
            best_i = i;
            best_j = rj;
            int32_t off = rj > i ? rj - i : i - rj;
            if (off > max_off) max_off = off;
        } else if (zdrop > 0) {
            int32_t di = i - best_i, dj = rj - best_j;
            if (di > dj) {
                if (best - rmax - (di - dj) * e_del > zdrop) break;
            } else if (best - rmax - (dj - di) * e_ins > zdrop) {
                break;
            }
        }
        int32_t k = beg;
        while (k < end && H[k] == 0 && E[k] == 0) ++k;
        beg = k;
        k = end;
        while (k >= beg && H[k] == 0 && E[k] == 0) --k;
        end = k + 2 < qlen ? k + 2 : qlen;
    }
    free(prof);
    free(H);
    free(E);
    if (cells) *cells += ncell;
    orc_ext_result r = {best, best_j + 1, best_i + 1, gbest_i + 1, gscore, max_off};
    return r;
}

/* crates/bwa-extend/src/sw.rs:806-824 (SuboptimalTracker::push_row). */
void orc_tracker_push_row(orc_tracker *b, int32_t row, int32_t row_max, int32_t minsc) {
    if (row_max < minsc) return;
    if (b->n > 0 && b->row[b->n - 1] + 1 == row) {
        if (b->score[b->n - 1] < row_max) {
            b->score[b->n - 1] = row_max;
            b->row[b->n - 1] = row;
        }
        return;
    }
    b->score[b->n] = row_max;
    b->row[b->n] = row;
    b->n++;
}

/* crates/bwa-extend/src/sw.rs:832-851 (SuboptimalTracker::finish). */
void orc_tracker_finish(const orc_tracker *b, int32_t gmax, int32_t te, int32_t max_sc,
                        int32_t *score2, int32_t *te2) {
    int32_t s2 = -1, t2 = -1;
    if (gmax > 0 && b->n > 0) {
        int32_t half = (gmax + max_sc - 1) / max_sc;
        int32_t lo = te - half, hi = te + half;
        for (int32_t k = 0; k < b->n; ++k) {
            if ((b->row[k] < lo || b->row[k] > hi) && b->score[k] > s2) {
                s2 = b->score[k];
                t2 = b->row[k];
            }
        }
    }
    *score2 = s2;
    *te2 = t2;
}

/* crates/bwa-neon/src/matesw.rs:676-686 with the lane rule of crates/bwa-mem/src/pe.rs:420-424:
 * 16 lanes when qlen*max_sc < 250, else 8. */
int32_t orc_lanes_for(int32_t qlen, int32_t max_sc) { return qlen * max_sc < 250 ? 16 : 8; }
int32_t orc_padded_qlen(int32_t qlen, int32_t max_sc) {
    int32_t l = orc_lanes_for(qlen, max_sc);
    return (qlen + l - 1) / l * l;
}

/* crates/bwa-extend/src/sw.rs:921-1075 (ksw_local_fwd). Gaps open from H; the query is padded
 * with score-0 columns to ceil(qlen/lanes)*lanes. */
orc_fwd_result orc_ksw_local_fwd(int32_t qlen_real, const uint8_t *query, int32_t tlen,
                                 const uint8_t *target, int32_t m, const int8_t *mat,
                                 int32_t o_del, int32_t e_del, int32_t o_ins, int32_t e_ins,
                                 int32_t minsc, int32_t endsc, int32_t max_sc, int32_t lanes,
                                 int64_t *cells) {
    const int32_t slen = (qlen_real + lanes - 1) / lanes;
    const int32_t qlen = slen * lanes;
    const int32_t oe_del = o_del + e_del, oe_ins = o_ins + e_ins;
    int32_t *hp = (int32_t *)calloc((size_t)qlen + 1, sizeof(int32_t));
    int32_t *hc = (int32_t *)calloc((size_t)qlen + 1, sizeof(int32_t));
    int32_t *e = (int32_t *)calloc((size_t)qlen + 1, sizeof(int32_t));
    int32_t *hsnap = (int32_t *)calloc((size_t)qlen + 1, sizeof(int32_t));
    orc_tracker b;
    b.score = (int32_t *)malloc(((size_t)tlen + 1) * sizeof(int32_t));
    b.row = (int32_t *)malloc(((size_t)tlen + 1) * sizeof(int32_t));
    b.n = 0;
    int32_t gmax = 0, te = -1;
    int64_t ncell = 0;
    for (int32_t i = 0; i < tlen; ++i) {
        const int8_t *srow = mat + (size_t)target[i] * m;
        int32_t f = 0, hdiag = 0, imax = 0;
        for (int32_t j = 0; j < qlen; ++j) {
            int32_t sc = j < qlen_real ? srow[query[j]] : 0;
            int32_t h = hdiag + sc;
            if (h < 0) h = 0;
            if (e[j] > h) h = e[j];
            if (f > h) h = f;
            if (h > imax) imax = h;
            hc[j] = h;
            int32_t en = e[j] - e_del, od = h - oe_del;
            if (od > en) en = od;
            e[j] = en > 0 ? en : 0;
            int32_t fn = f - e_ins, oi = h - oe_ins;
            if (oi > fn) fn = oi;
            f = fn > 0 ? fn : 0;
            hdiag = hp[j];
        }
        ncell += qlen;
        orc_tracker_push_row(&b, i, imax, minsc);
        if (imax > gmax) {
            gmax = imax;
            te = i;
            memcpy(hsnap, hc, (size_t)qlen * sizeof(int32_t));
            if (gmax >= endsc) break;
        }
        int32_t *tmp = hp;
        hp = hc;
        hc = tmp;
    }
    /* qe: smallest column holding the row maximum in the snapshot of row te (sw.rs:1052-1070);
     * scanning in column order with a strict '>' yields the smallest such column. */
    int32_t qe = -1;
    if (te >= 0) {
        int32_t bh = -1;
        for (int32_t j = 0; j < qlen; ++j)
            if (hsnap[j] > bh) { bh = hsnap[j]; qe = j; }
    }
    orc_fwd_result r;
    r.score = gmax;
    r.te = te;
    r.qe = qe;
    orc_tracker_finish(&b, gmax, te, max_sc, &r.score2, &r.te2);
    free(hp);
    free(hc);
    free(e);
    free(hsnap);
    free(b.score);
    free(b.row);
    if (cells) *cells += ncell;
    return r;
}

/* crates/bwa-extend/src/sw.rs:1108-1185 (ksw_align2). */
orc_ksw_result orc_ksw_align2(int32_t qlen, const uint8_t *query, int32_t tlen,
                              const uint8_t *target, int32_t m, const int8_t *mat, int32_t o_del,
                              int32_t e_del, int32_t o_ins, int32_t e_ins, int32_t minsc,
                              int32_t max_sc, int32_t lanes, int64_t *cells) {
    orc_fwd_result f = orc_ksw_local_fwd(qlen, query, tlen, target, m, mat, o_del, e_del, o_ins,
                                         e_ins, minsc, INT32_MAX, max_sc, lanes, cells);
    orc_ksw_result r = {f.score, -1, f.qe, -1, f.te, f.score2, f.te2};
    if (f.score < minsc || f.qe < 0) return r;
    /* the query prefix [0, qe] reversed; the whole target with its first te+1 bases reversed in
     * place, tail kept (sw.rs:1159-1161).  qe is always a real column: a padding cell of the first
     * row that reaches gmax would need a predecessor already at gmax.  The min() only restates
     * that bound. */
    int32_t rq = f.qe + 1 < qlen ? f.qe + 1 : qlen;
    uint8_t *qr = (uint8_t *)malloc((size_t)rq + 1);
    uint8_t *tr = (uint8_t *)malloc((size_t)tlen + 1);
    for (int32_t k = 0; k < rq; ++k) qr[k] = query[rq - 1 - k];
    memcpy(tr, target, (size_t)tlen);
    for (int32_t k = 0; k <= f.te; ++k) tr[k] = target[f.te - k];
    orc_fwd_result rv = orc_ksw_local_fwd(rq, qr, tlen, tr, m, mat, o_del, e_del, o_ins, e_ins,
                                          INT32_MAX, f.score, max_sc, lanes, cells);
    if (rv.score == f.score) {
        r.tb = f.te - rv.te;
        r.qb = f.qe - rv.qe;
    }
    free(qr);
    free(tr);
    return r;
}

/* crates/bwa-neon/src/matesw.rs:299-497 (batched_ksw_align2), one job: forward pass padded by the
 * query's own lane rule; reverse pass on reversed query[..=qe] and reversed target[..=te] only,
 * padded by the reversed query's own lane rule, stopping at the forward score. */
orc_ksw_result orc_batched_align2_job(int32_t qlen, const uint8_t *query, int32_t tlen,
                                      const uint8_t *target, int32_t m, const int8_t *mat,
                                      int32_t o_del, int32_t e_del, int32_t o_ins, int32_t e_ins,
                                      int32_t minsc, int32_t max_sc, int64_t *cells) {
    orc_ksw_result r = {0, -1, -1, -1, -1, -1, -1};
    if (qlen == 0 || tlen == 0) {
        /* an empty side gives no DP work and the default result (matesw.rs:964-966) */
        return r;
    }
    orc_fwd_result f =
        orc_ksw_local_fwd(qlen, query, tlen, target, m, mat, o_del, e_del, o_ins, e_ins, minsc,
                          INT32_MAX, max_sc, orc_lanes_for(qlen, max_sc), cells);
    r.score = f.score;
    r.qe = f.te >= 0 ? f.qe : -1;
    r.te = f.te;
    r.score2 = f.score2;
    r.te2 = f.te2;
    if (!(r.score >= minsc && r.qe >= 0)) return r;
    int32_t rq = r.qe + 1 < qlen ? r.qe + 1 : qlen;
    int32_t rt = r.te + 1;
    uint8_t *qr = (uint8_t *)malloc((size_t)rq + 1);
    uint8_t *tr = (uint8_t *)malloc((size_t)rt + 1);
    for (int32_t k = 0; k < rq; ++k) qr[k] = query[rq - 1 - k];
    for (int32_t k = 0; k < rt; ++k) tr[k] = target[rt - 1 - k];
    orc_fwd_result rv = orc_ksw_local_fwd(rq, qr, rt, tr, m, mat, o_del, e_del, o_ins, e_ins,
                                          INT32_MAX, r.score, max_sc, orc_lanes_for(rq, max_sc),
                                          cells);
    if (rv.score == r.score) {
        r.tb = r.te - rv.te;
        r.qb = r.qe - rv.qe;
    }
    free(qr);
    free(tr);
    return r;
}

/* ---------------------------------------------------------------------------------------------
 * Threaded batch drivers (static partition, one pthread per slice).
 * ------------------------------------------------------------------------------------------- */
typedef struct {
    int kind; /* 0 extension, 1 rescue */
    size_t lo, hi;
    const uint8_t *arena;
    const uint64_t *q_off, *t_off;
    const uint32_t *q_len, *t_len;
    const int32_t *h0;
    int32_t m;
    const int8_t *mat;
    int32_t o_del, e_del, o_ins, e_ins, w, end_bonus, zdrop, minsc, max_sc, batched;
    void *out;
    int64_t cells;
    int32_t *job_cells; /* optional: cells evaluated per job */
} orc_slice;

static void *orc_slice_run(void *arg) {
    orc_slice *s = (orc_slice *)arg;
    for (size_t k = s->lo; k < s->hi; ++k) {
        const uint8_t *q = s->arena + s->q_off[k], *t = s->arena + s->t_off[k];
        const int64_t c0 = s->cells;
        if (s->kind == 0) {
            ((orc_ext_result *)s->out)[k] = orc_ksw_extend2(
                (int32_t)s->q_len[k], q, (int32_t)s->t_len[k], t, s->m, s->mat, s->o_del,
                s->e_del, s->o_ins, s->e_ins, s->w, s->end_bonus, s->zdrop, s->h0[k], &s->cells);
        } else if (s->batched) {
            ((orc_ksw_result *)s->out)[k] = orc_batched_align2_job(
                (int32_t)s->q_len[k], q, (int32_t)s->t_len[k], t, s->m, s->mat, s->o_del,
                s->e_del, s->o_ins, s->e_ins, s->minsc, s->max_sc, &s->cells);
        } else {
            int32_t ql = (int32_t)s->q_len[k];
            ((orc_ksw_result *)s->out)[k] = orc_ksw_align2(
                ql, q, (int32_t)s->t_len[k], t, s->m, s->mat, s->o_del, s->e_del, s->o_ins,
                s->e_ins, s->minsc, s->max_sc, orc_lanes_for(ql, s->max_sc), &s->cells);
        }
        if (s->job_cells) s->job_cells[k] = (int32_t)(s->cells - c0);
    }
    return NULL;
}

static void orc_run_slices(orc_slice *proto, size_t n, int nthreads, int64_t *cells_out) {
    if (nthreads < 1) nthreads = 1;
    if ((size_t)nthreads > n && n > 0) nthreads = (int)n;
    orc_slice *sl = (orc_slice *)calloc((size_t)nthreads, sizeof(orc_slice));
    pthread_t *th = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
    for (int t = 0; t < nthreads; ++t) {
        sl[t] = *proto;
        sl[t].lo = n * (size_t)t / (size_t)nthreads;
        sl[t].hi = n * (size_t)(t + 1) / (size_t)nthreads;
        sl[t].cells = 0;
        if (nthreads == 1) orc_slice_run(&sl[t]);
        else pthread_create(&th[t], NULL, orc_slice_run, &sl[t]);
    }
    int64_t cells = 0;
    for (int t = 0; t < nthreads; ++t) {
        if (nthreads > 1) pthread_join(th[t], NULL);
        cells += sl[t].cells;
    }
    if (cells_out) *cells_out = cells;
    free(sl);
    free(th);
}

void orc_extend_batch(size_t n, const uint8_t *arena, const uint64_t *q_off, const uint32_t *q_len,
                      const uint64_t *t_off, const uint32_t *t_len, const int32_t *h0, int32_t m,
                      const int8_t *mat, int32_t o_del, int32_t e_del, int32_t o_ins,
                      int32_t e_ins, int32_t w, int32_t end_bonus, int32_t zdrop,
                      orc_ext_result *out, int64_t *cells_out, int nthreads) {
    orc_slice p;
    memset(&p, 0, sizeof p);
    p.kind = 0;
    p.arena = arena;
    p.q_off = q_off;
    p.q_len = q_len;
    p.t_off = t_off;
    p.t_len = t_len;
    p.h0 = h0;
    p.m = m;
    p.mat = mat;
    p.o_del = o_del;
    p.e_del = e_del;
    p.o_ins = o_ins;
    p.e_ins = e_ins;
    p.w = w;
    p.end_bonus = end_bonus;
    p.zdrop = zdrop;
    p.out = out;
    orc_run_slices(&p, n, nthreads, cells_out);
}

void orc_extend_batch_cells(size_t n, const uint8_t *arena, const uint64_t *q_off, const uint32_t *q_len,
                            const uint64_t *t_off, const uint32_t *t_len, const int32_t *h0, int32_t m,
                            const int8_t *mat, int32_t o_del, int32_t e_del, int32_t o_ins, int32_t e_ins,
                            int32_t w, int32_t end_bonus, int32_t zdrop, orc_ext_result *out,
                            int32_t *job_cells, int nthreads) {
    orc_slice p;
    memset(&p, 0, sizeof p);
    p.kind = 0;
    p.arena = arena;
    p.q_off = q_off;
    p.q_len = q_len;
    p.t_off = t_off;
    p.t_len = t_len;
    p.h0 = h0;
    p.m = m;
    p.mat = mat;
    p.o_del = o_del;
    p.e_del = e_del;
    p.o_ins = o_ins;
    p.e_ins = e_ins;
    p.w = w;
    p.end_bonus = end_bonus;
    p.zdrop = zdrop;
    p.out = out;
    p.job_cells = job_cells;
    orc_run_slices(&p, n, nthreads, NULL);
}

void orc_align2_batch(size_t n, const uint8_t *arena, const uint64_t *q_off, const uint32_t *q_len,
                      const uint64_t *t_off, const uint32_t *t_len, int32_t m, const int8_t *mat,
                      int32_t o_del, int32_t e_del, int32_t o_ins, int32_t e_ins, int32_t minsc,
                      int32_t max_sc, int batched_semantics, orc_ksw_result *out,
                      int64_t *cells_out, int nthreads) {
    orc_slice p;
    memset(&p, 0, sizeof p);
    p.kind = 1;
    p.arena = arena;
    p.q_off = q_off;
    p.q_len = q_len;
    p.t_off = t_off;
    p.t_len = t_len;
    p.m = m;
    p.mat = mat;
    p.o_del = o_del;
    p.e_del = e_del;
    p.o_ins = o_ins;
    p.e_ins = e_ins;
    p.minsc = minsc;
    p.max_sc = max_sc;
    p.batched = batched_semantics;
    p.out = out;
    orc_run_slices(&p, n, nthreads, cells_out);
}
