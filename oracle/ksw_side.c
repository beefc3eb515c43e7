/*
 * CPU oracle (test infrastructure only): bwa's band doubling around ksw_extend2 for one side of
 * the seed extensions, in the two forms the reference has.
 *
 *   orc_extend_side  crates/bwa-mem/src/lib.rs:332-373     (extend_side, one job's band tries)
 *   orc_run_side     crates/bwa-mem/src/across.rs:1357-1460 (run_side, rounds over a batch:
 *                    gather the active jobs, extend them all at band w << round, accept or requeue)
 *
 * Acceptance (lib.rs:360, across.rs:1444-1445, from bwamem.cpp:2495-2496): the score equals the
 * previous band's (prev before the first), or max_off < (w >> 1) + (w >> 2), or the last try.
 */
#include <stdlib.h>
#include <string.h>

#include "ksw_oracle.h"

static int orc_side_accept(const orc_ext_result *r, int32_t prev, int32_t w, int32_t round,
                           int32_t max_band_try) {
    return r->score == prev || r->max_off < (w >> 1) + (w >> 2) || round + 1 == max_band_try;
}

orc_ext_result orc_extend_side(int32_t qlen, const uint8_t *query, int32_t tlen, const uint8_t *target,
                               int32_t m, const int8_t *mat, int32_t o_del, int32_t e_del, int32_t o_ins,
                               int32_t e_ins, int32_t w, int32_t end_bonus, int32_t zdrop, int32_t h0,
                               int32_t prev0, int32_t max_band_try, int32_t *w_out) {
    int32_t prev = prev0;
    for (int32_t band_try = 0;; ++band_try) {
        const int32_t wb = w << band_try;
        orc_ext_result r = orc_ksw_extend2(qlen, query, tlen, target, m, mat, o_del, e_del, o_ins, e_ins,
                                           wb, end_bonus, zdrop, h0, NULL);
        if (orc_side_accept(&r, prev, wb, band_try, max_band_try)) {
            *w_out = wb;
            return r;
        }
        prev = r.score;
    }
}

void orc_run_side(size_t n, const uint8_t *arena, const uint64_t *q_off, const uint32_t *q_len,
                  const uint64_t *t_off, const uint32_t *t_len, const int32_t *h0, const int32_t *prev0,
                  int32_t m, const int8_t *mat, int32_t o_del, int32_t e_del, int32_t o_ins, int32_t e_ins,
                  int32_t w, int32_t end_bonus, int32_t zdrop, int32_t max_band_try, orc_ext_result *out,
                  int32_t *w_out, int nthreads) {
    int32_t *prev = (int32_t *)malloc((n ? n : 1) * sizeof(int32_t));
    uint8_t *active = (uint8_t *)malloc(n ? n : 1);
    size_t *idx = (size_t *)malloc((n ? n : 1) * sizeof(size_t));
    uint64_t *qo = (uint64_t *)malloc((n ? n : 1) * 8), *to = (uint64_t *)malloc((n ? n : 1) * 8);
    uint32_t *ql = (uint32_t *)malloc((n ? n : 1) * 4), *tl = (uint32_t *)malloc((n ? n : 1) * 4);
    int32_t *hh = (int32_t *)malloc((n ? n : 1) * 4);
    orc_ext_result *res = (orc_ext_result *)malloc((n ? n : 1) * sizeof(orc_ext_result));
    memcpy(prev, prev0, n * sizeof(int32_t));
    memset(active, 1, n);
    for (int32_t round = 0; round < max_band_try; ++round) {
        /* step 1: the still-active jobs */
        size_t na = 0;
        for (size_t k = 0; k < n; ++k)
            if (active[k]) {
                idx[na] = k;
                qo[na] = q_off[k];
                to[na] = t_off[k];
                ql[na] = q_len[k];
                tl[na] = t_len[k];
                hh[na] = h0[k];
                ++na;
            }
        if (na == 0) break;
        /* step 2: the batch at this round's band */
        const int32_t wr = w << round;
        orc_extend_batch(na, arena, qo, ql, to, tl, hh, m, mat, o_del, e_del, o_ins, e_ins, wr, end_bonus,
                         zdrop, res, NULL, nthreads);
        /* step 3: accept (the result and its band are final) or requeue with prev = score */
        for (size_t a = 0; a < na; ++a) {
            const size_t k = idx[a];
            if (orc_side_accept(&res[a], prev[k], wr, round, max_band_try)) {
                active[k] = 0;
                out[k] = res[a];
                w_out[k] = wr;
            } else {
                prev[k] = res[a].score;
            }
        }
    }
    free(prev);
    free(active);
    free(idx);
    free(qo);
    free(to);
    free(ql);
    free(tl);
    free(hh);
    free(res);
}
