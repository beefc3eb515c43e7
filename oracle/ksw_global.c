/*
 * CPU oracle (test infrastructure only): banded global alignment with traceback.
 *
 *   orc_ksw_global2   crates/bwa-extend/src/sw.rs:561-740 (ksw_global2), with push_cigar
 *                     sw.rs:515-524.  CIGAR words are len << 4 | op, op 0 = M, 1 = I, 2 = D.
 *   orc_global_batch  one call per job over a batch (each job with its own band w).
 *   orc_brute_global  exhaustive enumeration of global alignments on tiny inputs under the same
 *                     gap rules (gaps open only from an aligned pair, a leading gap from the
 *                     origin), the check that pins the score.
 */
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "ksw_oracle.h"

#define ORC_MINUS_INF (-0x40000000)

static void push_cigar(uint32_t *cig, int32_t *n, uint32_t op, uint32_t len) {
    if (*n > 0 && (cig[*n - 1] & 0xf) == op) {
        cig[*n - 1] += len << 4;
        return;
    }
    cig[(*n)++] = (len << 4) | op;
}

int32_t orc_ksw_global2(int32_t qlen, const uint8_t *query, int32_t tlen, const uint8_t *target, int32_t m,
                        const int8_t *mat, int32_t o_del, int32_t e_del, int32_t o_ins, int32_t e_ins, int32_t w,
                        uint32_t *cigar, int32_t *n_cigar) {
    const int32_t oe_del = o_del + e_del, oe_ins = o_ins + e_ins;
    if (w < 0) w = 0;
    const int32_t n_col = qlen < 2 * w + 1 ? qlen : 2 * w + 1;
    uint8_t *z = (uint8_t *)calloc((size_t)n_col * (size_t)tlen + 1, 1);
    int8_t *qp = (int8_t *)malloc((size_t)qlen * m + 1);
    for (int32_t c = 0; c < m; ++c)
        for (int32_t j = 0; j < qlen; ++j) qp[c * qlen + j] = mat[c * m + query[j]];
    int32_t *eh_h = (int32_t *)malloc(((size_t)qlen + 1) * sizeof(int32_t));
    int32_t *eh_e = (int32_t *)malloc(((size_t)qlen + 1) * sizeof(int32_t));
    for (int32_t j = 0; j <= qlen; ++j) eh_h[j] = eh_e[j] = ORC_MINUS_INF;
    eh_h[0] = 0;
    for (int32_t j = 1; j <= qlen && j <= w; ++j) eh_h[j] = -(o_ins + e_ins * j);
    for (int32_t i = 0; i < tlen; ++i) {
        int32_t f = ORC_MINUS_INF;
        const int32_t beg = i > w ? i - w : 0;
        const int32_t end = i + w + 1 < qlen ? i + w + 1 : qlen;
        int32_t h1 = beg == 0 ? -(o_del + e_del * (i + 1)) : ORC_MINUS_INF;
        const int8_t *q = qp + (size_t)target[i] * qlen;
        uint8_t *zr = z + (size_t)i * n_col;
        for (int32_t j = beg; j < end; ++j) {
            int32_t mm = eh_h[j], e = eh_e[j];
            eh_h[j] = h1;
            mm += q[j];
            uint8_t d = mm < e ? 1 : 0;  /* ties: M over E over F */
            int32_t h = mm >= e ? mm : e;
            d = h >= f ? d : 2;
            h = h >= f ? h : f;
            h1 = h;
            int32_t t = mm - oe_del;
            e -= e_del;
            d |= e > t ? 1 << 2 : 0;
            e = e > t ? e : t;
            eh_e[j] = e;
            t = mm - oe_ins;
            f -= e_ins;
            d |= f > t ? 2 << 4 : 0;
            f = f > t ? f : t;
            zr[j - beg] = d;
        }
        eh_h[end] = h1;
        eh_e[end] = ORC_MINUS_INF;
    }
    const int32_t score = eh_h[qlen];
    int32_t n = 0;
    int64_t i = (int64_t)tlen - 1;
    int64_t k = ((int64_t)tlen - 1 + w + 1 < qlen ? (int64_t)tlen - 1 + w + 1 : qlen) - 1;
    int which = 0;
    while (i >= 0 && k >= 0) {
        const int64_t beg = i > w ? i - w : 0;
        const uint8_t d = z[i * n_col + (k - beg)];
        which = (d >> (which << 1)) & 3;
        if (which == 0) {
            push_cigar(cigar, &n, 0, 1);
            --i;
            --k;
        } else if (which == 1) {
            push_cigar(cigar, &n, 2, 1);
            --i;
        } else {
            push_cigar(cigar, &n, 1, 1);
            --k;
        }
    }
    if (i >= 0) push_cigar(cigar, &n, 2, (uint32_t)(i + 1));
    if (k >= 0) push_cigar(cigar, &n, 1, (uint32_t)(k + 1));
    for (int32_t a = 0, b = n - 1; a < b; ++a, --b) {
        const uint32_t t = cigar[a];
        cigar[a] = cigar[b];
        cigar[b] = t;
    }
    *n_cigar = n;
    free(z);
    free(qp);
    free(eh_h);
    free(eh_e);
    return score;
}

typedef struct {
    size_t lo, hi;
    const uint8_t *arena;
    const uint64_t *q_off, *t_off, *cig_off;
    const uint32_t *q_len, *t_len;
    const int32_t *w;
    const int8_t *mat;
    int32_t o_del, e_del, o_ins, e_ins;
    int32_t *score, *n_cigar;
    uint32_t *cigar;
} orc_glb_slice;

static void *orc_glb_run(void *arg) {
    orc_glb_slice *s = (orc_glb_slice *)arg;
    for (size_t k = s->lo; k < s->hi; ++k)
        s->score[k] = orc_ksw_global2((int32_t)s->q_len[k], s->arena + s->q_off[k], (int32_t)s->t_len[k],
                                      s->arena + s->t_off[k], 5, s->mat, s->o_del, s->e_del, s->o_ins, s->e_ins,
                                      s->w[k], s->cigar + s->cig_off[k], &s->n_cigar[k]);
    return NULL;
}

/* cigar receives job k's runs at cig_off[k]; a job needs at most qlen + tlen + 1 words. */
void orc_global_batch(size_t n, const uint8_t *arena, const uint64_t *q_off, const uint32_t *q_len,
                      const uint64_t *t_off, const uint32_t *t_len, const int32_t *w, const int8_t *mat,
                      int32_t o_del, int32_t e_del, int32_t o_ins, int32_t e_ins, int32_t *score,
                      const uint64_t *cig_off, uint32_t *cigar, int32_t *n_cigar, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if ((size_t)nthreads > n && n > 0) nthreads = (int)n;
    orc_glb_slice *sl = (orc_glb_slice *)calloc((size_t)nthreads, sizeof(orc_glb_slice));
    pthread_t *th = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
    for (int t = 0; t < nthreads; ++t) {
        orc_glb_slice *s = &sl[t];
        s->lo = n * (size_t)t / (size_t)nthreads;
        s->hi = n * (size_t)(t + 1) / (size_t)nthreads;
        s->arena = arena;
        s->q_off = q_off;
        s->t_off = t_off;
        s->cig_off = cig_off;
        s->q_len = q_len;
        s->t_len = t_len;
        s->w = w;
        s->mat = mat;
        s->o_del = o_del;
        s->e_del = e_del;
        s->o_ins = o_ins;
        s->e_ins = e_ins;
        s->score = score;
        s->n_cigar = n_cigar;
        s->cigar = cigar;
        if (nthreads == 1) orc_glb_run(s);
        else pthread_create(&th[t], NULL, orc_glb_run, s);
    }
    if (nthreads > 1)
        for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    free(sl);
    free(th);
}

/* Exhaustive search over global alignments of tiny sequences: the best score of a path that
 * aligns all of query against all of target, where an interior gap run is entered only from an
 * aligned pair, a leading gap run only from the origin, and a deletion and an insertion are never
 * adjacent.  state: 0 = last step an aligned pair (or the origin), 1 = inside a deletion, 2 =
 * inside an insertion. */
static int32_t brute_rec(int32_t i, int32_t j, int state, int32_t qlen, const uint8_t *q,
                         int32_t tlen, const uint8_t *t, const int8_t *mat, int32_t o_del, int32_t e_del,
                         int32_t o_ins, int32_t e_ins) {
    if (i == tlen && j == qlen) return 0;
    int32_t best = ORC_MINUS_INF;
    if (i < tlen && j < qlen) {
        const int32_t v = brute_rec(i + 1, j + 1, 0, qlen, q, tlen, t, mat, o_del, e_del, o_ins, e_ins);
        if (v > ORC_MINUS_INF) best = v + mat[t[i] * 5 + q[j]] > best ? v + mat[t[i] * 5 + q[j]] : best;
    }
    if (i < tlen && state != 2) {  /* deletion: consumes a target base */
        const int32_t cost = state == 1 ? e_del : o_del + e_del;
        const int32_t v = brute_rec(i + 1, j, 1, qlen, q, tlen, t, mat, o_del, e_del, o_ins, e_ins);
        if (v > ORC_MINUS_INF && v - cost > best) best = v - cost;
    }
    if (j < qlen && state != 1) {  /* insertion: consumes a query base */
        const int32_t cost = state == 2 ? e_ins : o_ins + e_ins;
        const int32_t v = brute_rec(i, j + 1, 2, qlen, q, tlen, t, mat, o_del, e_del, o_ins, e_ins);
        if (v > ORC_MINUS_INF && v - cost > best) best = v - cost;
    }
    return best;
}

int32_t orc_brute_global(int32_t qlen, const uint8_t *query, int32_t tlen, const uint8_t *target, const int8_t *mat,
                         int32_t o_del, int32_t e_del, int32_t o_ins, int32_t e_ins) {
    return brute_rec(0, 0, 0, qlen, query, tlen, target, mat, o_del, e_del, o_ins, e_ins);
}
