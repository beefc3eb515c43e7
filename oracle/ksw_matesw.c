/*
 * CPU oracle (test infrastructure only): one round of the reference's batched mate rescue, from an
 * anchor region and the mate's sequence to the rescue windows and their batched_ksw_align2 results.
 *
 *   orc_infer_dir       crates/bwa-mem/src/pe.rs:192-209        (mem_infer_dir)
 *   orc_depos           crates/bwa-index/src/bntseq.rs:237-243  (depos)
 *   orc_pos2rid         crates/bwa-index/src/bntseq.rs:337-367  (pos2rid)
 *   orc_fetch_window    crates/bwa-mem/src/pe.rs:211-269        (bns_fetch_seq, coordinates)
 *   orc_ref_base        crates/bwa-index/src/fmindex.rs:281-310 (bases of the doubled space)
 *   orc_matesw_collect  crates/bwa-mem/src/pe.rs:596-679        (matesw_collect)
 *   orc_matesw_batch    crates/bwa-mem/src/pe.rs:1040-1113      (one round: collect, flatten,
 *                       batched_ksw_align2 per job as orc_batched_align2_job)
 *
 * The reference is given as its forward strand, one code per byte; position p >= l_pac of the
 * doubled space is the complement of forward position 2*l_pac - 1 - p.
 */
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "ksw_oracle.h"

void orc_infer_dir(int64_t l_pac, int64_t b1, int64_t b2, int32_t *dir, int64_t *dist) {
    const int r1 = b1 >= l_pac, r2 = b2 >= l_pac;
    const int64_t p2 = r1 == r2 ? b2 : (l_pac << 1) - 1 - b2;
    *dist = p2 > b1 ? p2 - b1 : b1 - p2;
    *dir = (r1 == r2 ? 0 : 1) ^ (p2 > b1 ? 0 : 3);
}

int64_t orc_depos(int64_t l_pac, int64_t pos, int *is_rev) {
    *is_rev = pos >= l_pac;
    return pos >= l_pac ? (l_pac << 1) - 1 - pos : pos;
}

int32_t orc_pos2rid(int64_t l_pac, const int64_t *off, int32_t n, int64_t pos_f) {
    if (pos_f >= l_pac) return -1;
    int32_t left = 0, mid = 0, right = n;
    while (left < right) {
        mid = (left + right) >> 1;
        if (pos_f >= off[mid]) {
            if (mid == n - 1) break;
            if (pos_f < off[mid + 1]) break;
            left = mid + 1;
        } else {
            right = mid;
        }
    }
    return mid;
}

void orc_fetch_window(int64_t l_pac, const int64_t *off, const int32_t *len, int32_t n, int64_t rb, int64_t mid,
                      int64_t re, int64_t *rb_out, int64_t *re_out, int32_t *rid_out) {
    if (re < rb) {
        const int64_t t = rb;
        rb = re;
        re = t;
    }
    int is_rev;
    const int64_t fwd_mid = orc_depos(l_pac, mid, &is_rev);
    const int32_t rid = orc_pos2rid(l_pac, off, n, fwd_mid);
    int64_t far_beg = 0, far_end = l_pac << 1;
    if (rid >= 0) {
        far_beg = off[rid];
        far_end = off[rid] + len[rid];
    }
    if (is_rev) {
        const int64_t ob = far_beg;
        far_beg = (l_pac << 1) - far_end;
        far_end = (l_pac << 1) - ob;
    }
    *rb_out = rb > far_beg ? rb : far_beg;
    *re_out = re < far_end ? re : far_end;
    *rid_out = rid;
}

uint8_t orc_ref_base(const uint8_t *fwd, int64_t l_pac, int64_t p) {
    if (p < l_pac) return fwd[p];
    const uint8_t c = fwd[(l_pac << 1) - 1 - p];
    return c < 4 ? (uint8_t)(3 - c) : c;
}

/* Windows of one anchor (flags as in bwg.h: 1 valid, 2 reverse-complemented mate, 4 skipped). */
void orc_matesw_collect(int64_t l_pac, const int64_t *off, const int32_t *len, int32_t n_contigs,
                        const int32_t *low, const int32_t *high, const int32_t *failed, int32_t min_seed_len,
                        int64_t a_rb, int32_t a_rid, int64_t l_ms, const int64_t *ma_rb, size_t n_ma,
                        orc_rescue_window *win) {
    int skip[4];
    for (int r = 0; r < 4; ++r) skip[r] = failed[r] ? 1 : 0;
    for (size_t e = 0; e < n_ma; ++e) {
        int32_t r;
        int64_t dist;
        orc_infer_dir(l_pac, a_rb, ma_rb[e], &r, &dist);
        if (dist >= low[r] && dist <= high[r]) skip[r] = 1;
    }
    const int none = skip[0] + skip[1] + skip[2] + skip[3] == 4;
    for (int r = 0; r < 4; ++r) {
        orc_rescue_window *w = &win[r];
        const int is_rev = (r >> 1) != (r & 1), is_larger = (r >> 1) == 0;
        w->rb = w->re = 0;
        w->rid = -1;
        w->flags = (skip[r] ? 4 : 0) | (is_rev ? 2 : 0);
        if (skip[r] || none) continue;
        const int64_t lo = low[r], hi = high[r];
        int64_t rb0, re0;
        if (!is_rev) {
            rb0 = is_larger ? a_rb + lo : a_rb - hi;
            re0 = (is_larger ? a_rb + hi : a_rb - lo) + l_ms;
        } else {
            rb0 = (is_larger ? a_rb + lo : a_rb - hi) - l_ms;
            re0 = is_larger ? a_rb + hi : a_rb - lo;
        }
        if (rb0 < 0) rb0 = 0;
        if (re0 > (l_pac << 1)) re0 = l_pac << 1;
        if (rb0 < re0) {
            int64_t rb, re;
            int32_t rid;
            orc_fetch_window(l_pac, off, len, n_contigs, rb0, (rb0 + re0) >> 1, re0, &rb, &re, &rid);
            w->rb = rb;
            w->re = re;
            w->rid = rid;
            if (a_rid == rid && re - rb >= min_seed_len) w->flags |= 1;
        }
    }
}

typedef struct {
    size_t lo, hi;
    int64_t l_pac;
    const uint8_t *fwd;
    const int64_t *off;
    const int32_t *len;
    int32_t n_contigs;
    const int32_t *low, *high, *failed;
    int32_t min_seed_len;
    const orc_mate_anchor *anchors;
    const int64_t *ma_rb;
    const uint8_t *arena;
    const int8_t *mat;
    int32_t o_del, e_del, o_ins, e_ins, minsc, max_sc;
    orc_rescue_window *win;
    orc_ksw_result *out;
} orc_msw_slice;

static void *orc_msw_run(void *arg) {
    orc_msw_slice *s = (orc_msw_slice *)arg;
    for (size_t k = s->lo; k < s->hi; ++k) {
        const orc_mate_anchor *a = &s->anchors[k];
        orc_matesw_collect(s->l_pac, s->off, s->len, s->n_contigs, s->low, s->high, s->failed, s->min_seed_len,
                           a->rb, a->rid, a->l_ms, s->ma_rb + a->ma_begin, (size_t)(a->ma_end - a->ma_begin),
                           &s->win[4 * k]);
        const uint8_t *ms = s->arena + a->ms_off;
        uint8_t *rc = (uint8_t *)malloc((size_t)a->l_ms + 1);
        for (uint32_t j = 0; j < a->l_ms; ++j) {
            const uint8_t c = ms[a->l_ms - 1 - j];
            rc[j] = c < 4 ? (uint8_t)(3 - c) : c;
        }
        for (int r = 0; r < 4; ++r) {
            const orc_rescue_window *w = &s->win[4 * k + r];
            orc_ksw_result res = {0, -1, -1, -1, -1, -1, -1};
            if (w->flags & 1) {
                const int64_t tl = w->re - w->rb;
                uint8_t *t = (uint8_t *)malloc((size_t)tl + 1);
                for (int64_t i = 0; i < tl; ++i) t[i] = orc_ref_base(s->fwd, s->l_pac, w->rb + i);
                res = orc_batched_align2_job((int32_t)a->l_ms, (w->flags & 2) ? rc : ms, (int32_t)tl, t, 5, s->mat,
                                             s->o_del, s->e_del, s->o_ins, s->e_ins, s->minsc, s->max_sc, NULL);
                free(t);
            }
            s->out[4 * k + r] = res;
        }
        free(rc);
    }
    return NULL;
}

void orc_matesw_batch(int64_t l_pac, const uint8_t *fwd, const int64_t *off, const int32_t *len, int32_t n_contigs,
                      const int32_t *low, const int32_t *high, const int32_t *failed, int32_t min_seed_len,
                      size_t n, const orc_mate_anchor *anchors, const int64_t *ma_rb, const uint8_t *arena,
                      const int8_t *mat, int32_t o_del, int32_t e_del, int32_t o_ins, int32_t e_ins, int32_t minsc,
                      int32_t max_sc, orc_rescue_window *win, orc_ksw_result *out, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if ((size_t)nthreads > n && n > 0) nthreads = (int)n;
    orc_msw_slice *sl = (orc_msw_slice *)calloc((size_t)nthreads, sizeof(orc_msw_slice));
    pthread_t *th = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
    for (int t = 0; t < nthreads; ++t) {
        orc_msw_slice *s = &sl[t];
        s->lo = n * (size_t)t / (size_t)nthreads;
        s->hi = n * (size_t)(t + 1) / (size_t)nthreads;
        s->l_pac = l_pac;
        s->fwd = fwd;
        s->off = off;
        s->len = len;
        s->n_contigs = n_contigs;
        s->low = low;
        s->high = high;
        s->failed = failed;
        s->min_seed_len = min_seed_len;
        s->anchors = anchors;
        s->ma_rb = ma_rb;
        s->arena = arena;
        s->mat = mat;
        s->o_del = o_del;
        s->e_del = e_del;
        s->o_ins = o_ins;
        s->e_ins = e_ins;
        s->minsc = minsc;
        s->max_sc = max_sc;
        s->win = win;
        s->out = out;
        if (nthreads == 1) orc_msw_run(s);
        else pthread_create(&th[t], NULL, orc_msw_run, s);
    }
    if (nthreads > 1)
        for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    free(sl);
    free(th);
}
