/*
 * Brute-force checkers for the CPU oracle (test infrastructure only; see ksw_oracle.c).
 *
 * These compute the same quantities as the oracle from first principles, with none of the
 * reference's band, early-exit or in-place array machinery, so that the oracle is pinned by
 * something other than itself:
 *
 *   orc_brute_extend_unbanded  the unbanded recurrence the reference's own test compares against
 *                              (crates/bwa-extend/src/sw.rs:1255-1305, used at :1326-1356)
 *   orc_brute_gotoh_local      textbook Gotoh local alignment, full H/E/F matrices
 *   orc_brute_enum_local       explicit enumeration of every local alignment path (tiny inputs)
 */
#include "ksw_oracle.h"

#include <stdlib.h>
#include <string.h>

int32_t orc_brute_extend_unbanded(int32_t qlen, const uint8_t *query, int32_t tlen,
                                  const uint8_t *target, int32_t m, const int8_t *mat,
                                  int32_t o_del, int32_t e_del, int32_t o_ins, int32_t e_ins,
                                  int32_t h0) {
    /* Full (tlen+1) x (qlen+1) matrices, row -1 / column -1 boundary explicit: H[r][c] is the
     * cell (r-1, c-1).  M opens gaps (extension convention). */
    const int32_t R = tlen + 1, C = qlen + 1;
    int32_t *H = calloc((size_t)R * C, 4), *E = calloc((size_t)R * C, 4), *F = calloc((size_t)R * C, 4);
    const int32_t oe_del = o_del + e_del, oe_ins = o_ins + e_ins;
    /* row -1: h0 at column -1, then the insertion ramp while it stays positive */
    H[0] = h0;
    if (qlen >= 1) H[1] = h0 > oe_ins ? h0 - oe_ins : 0;
    for (int32_t c = 2; c <= qlen && H[c - 1] > e_ins; ++c) H[c] = H[c - 1] - e_ins;
    int32_t best = h0;
    for (int32_t r = 1; r <= tlen; ++r) {
        int32_t left = h0 - (o_del + e_del * r);
        H[r * C] = left > 0 ? left : 0; /* column -1 reached by a deletion from the origin */
        for (int32_t c = 1; c <= qlen; ++c) {
            int32_t diag = H[(r - 1) * C + (c - 1)];
            int32_t mm = diag != 0 ? diag + mat[target[r - 1] * m + query[c - 1]] : 0;
            /* E(r,c) is the deletion score entering this cell, computed from the row above */
            int32_t e = E[(r - 1) * C + c]; /* carried value for this cell */
            int32_t f = F[r * C + (c - 1)];
            int32_t h = mm;
            if (e > h) h = e;
            if (f > h) h = f;
            H[r * C + c] = h;
            if (h > best) best = h;
            int32_t to = mm - oe_del > 0 ? mm - oe_del : 0;
            E[r * C + c] = e - e_del > to ? e - e_del : to;
            to = mm - oe_ins > 0 ? mm - oe_ins : 0;
            F[r * C + c] = f - e_ins > to ? f - e_ins : to;
        }
    }
    free(H);
    free(E);
    free(F);
    return best;
}

/* Scores for the padded query of the rescue kernel: code 5 (padding) scores 0 against anything. */
static int32_t pad_score(int32_t m, const int8_t *mat, int32_t t, int32_t q) {
    return q >= m ? 0 : mat[t * m + q];
}

int32_t orc_brute_gotoh_local(int32_t qlen, const uint8_t *query, int32_t tlen,
                              const uint8_t *target, int32_t m, const int8_t *mat, int32_t o_del,
                              int32_t e_del, int32_t o_ins, int32_t e_ins) {
    const int32_t NEG = -(1 << 28);
    const int32_t R = tlen + 1, C = qlen + 1;
    int32_t *H = calloc((size_t)R * C, 4), *E = malloc((size_t)R * C * 4), *F = malloc((size_t)R * C * 4);
    for (int32_t k = 0; k < R * C; ++k) E[k] = F[k] = NEG;
    int32_t best = 0;
    for (int32_t r = 1; r <= tlen; ++r)
        for (int32_t c = 1; c <= qlen; ++c) {
            int32_t up = r > 1 ? H[(r - 1) * C + c] : 0;
            int32_t lf = c > 1 ? H[r * C + c - 1] : 0;
            int32_t e = E[(r - 1) * C + c] - e_del, eo = up - (o_del + e_del);
            E[r * C + c] = e > eo ? e : eo;
            int32_t f = F[r * C + c - 1] - e_ins, fo = lf - (o_ins + e_ins);
            F[r * C + c] = f > fo ? f : fo;
            int32_t h = H[(r - 1) * C + c - 1] + pad_score(m, mat, target[r - 1], query[c - 1]);
            if (h < 0) h = 0;
            if (E[r * C + c] > h) h = E[r * C + c];
            if (F[r * C + c] > h) h = F[r * C + c];
            H[r * C + c] = h;
            if (h > best) best = h;
        }
    free(H);
    free(E);
    free(F);
    return best;
}

/* Explicit path enumeration.  A local alignment path starts with an aligned pair at (i0, j0) and
 * continues with steps M (i+1, j+1), D (i+1, j) or I (i, j+1); every maximal run of D costs
 * o_del + len*e_del, every maximal run of I costs o_ins + len*e_ins, and a D run may directly
 * follow an I run and vice versa (gaps open from H).  best[i][j] is the best score over all paths
 * ending at (i, j). */
typedef struct {
    int32_t qlen, tlen, m;
    const uint8_t *q, *t;
    const int8_t *mat;
    int32_t o_del, e_del, o_ins, e_ins;
    int32_t *best;
} enum_ctx;

static void enum_dfs(enum_ctx *c, int32_t i, int32_t j, int32_t score, int last /*0 M,1 D,2 I*/) {
    int32_t *b = &c->best[i * c->qlen + j];
    if (score > *b) *b = score;
    if (i + 1 < c->tlen && j + 1 < c->qlen)
        enum_dfs(c, i + 1, j + 1, score + pad_score(c->m, c->mat, c->t[i + 1], c->q[j + 1]), 0);
    if (i + 1 < c->tlen)
        enum_dfs(c, i + 1, j, score - (last == 1 ? c->e_del : c->o_del + c->e_del), 1);
    if (j + 1 < c->qlen)
        enum_dfs(c, i, j + 1, score - (last == 2 ? c->e_ins : c->o_ins + c->e_ins), 2);
}

void orc_brute_enum_local(int32_t qlen, const uint8_t *query, int32_t tlen, const uint8_t *target,
                          int32_t m, const int8_t *mat, int32_t o_del, int32_t e_del,
                          int32_t o_ins, int32_t e_ins, int32_t *score, int32_t *te, int32_t *qe) {
    enum_ctx c = {qlen, tlen, m, query, target, mat, o_del, e_del, o_ins, e_ins, NULL};
    c.best = calloc((size_t)(qlen > 0 ? qlen : 1) * (tlen > 0 ? tlen : 1), 4);
    for (int32_t i = 0; i < tlen; ++i)
        for (int32_t j = 0; j < qlen; ++j) enum_dfs(&c, i, j, pad_score(m, mat, target[i], query[j]), 0);
    int32_t g = 0, bi = -1, bj = -1;
    for (int32_t i = 0; i < tlen; ++i)
        for (int32_t j = 0; j < qlen; ++j)
            if (c.best[i * qlen + j] > g) g = c.best[i * qlen + j];
    if (g > 0) {
        for (int32_t i = 0; i < tlen && bi < 0; ++i)
            for (int32_t j = 0; j < qlen; ++j)
                if (c.best[i * qlen + j] == g) { bi = i; bj = j; break; }
    }
    *score = g;
    *te = bi;
    *qe = bj;
    free(c.best);
}

/* Explicit enumeration of every alignment prefix from the origin under bwa's extension and
 * global-alignment gap rules (crates/bwa-extend/src/sw.rs:364-414 and :597-636): a deletion (D,
 * consumes a target base) or insertion (I, consumes a query base) run opens only from an aligned
 * pair or from the origin, so a D run is never next to an I run; a run costs o + len*e.  Every cell
 * visited, the origin's boundary cells (-1, j) and (i, -1) included, must lie in the band
 * |i - j| <= w.  For each cell (i, j) of the tlen x qlen matrix, Mx/Ex/Fx receive the best score of
 * a prefix path ending there with an aligned pair / a D / an I as its last step (the origin scores
 * `origin`), or ORC_ENUM_NONE when no such path exists.  No recurrence is evaluated: every path is
 * walked.  Tiny inputs only (the number of paths grows like the Delannoy numbers). */
typedef struct {
    int32_t qlen, tlen, w;
    const uint8_t *q, *t;
    const int8_t *mat;
    int32_t o_del, e_del, o_ins, e_ins;
    int32_t *M, *E, *F;
} prefix_ctx;

static int in_band(const prefix_ctx *c, int32_t i, int32_t j) {
    const int32_t d = i - j;
    return d <= c->w && -d <= c->w;
}

static void prefix_dfs(prefix_ctx *c, int32_t i, int32_t j, int32_t score, int last /*0 M/origin,1 D,2 I*/) {
    if (i >= 0 && j >= 0) {
        int32_t *slot = (last == 0 ? c->M : last == 1 ? c->E : c->F) + i * c->qlen + j;
        if (*slot == ORC_ENUM_NONE || score > *slot) *slot = score;
    }
    if (i + 1 < c->tlen && j + 1 < c->qlen && in_band(c, i + 1, j + 1))
        prefix_dfs(c, i + 1, j + 1, score + c->mat[c->t[i + 1] * 5 + c->q[j + 1]], 0);
    if (last != 2 && i + 1 < c->tlen && in_band(c, i + 1, j))
        prefix_dfs(c, i + 1, j, score - (last == 1 ? c->e_del : c->o_del + c->e_del), 1);
    if (last != 1 && j + 1 < c->qlen && in_band(c, i, j + 1))
        prefix_dfs(c, i, j + 1, score - (last == 2 ? c->e_ins : c->o_ins + c->e_ins), 2);
}

void orc_enum_prefix(int32_t qlen, const uint8_t *query, int32_t tlen, const uint8_t *target, const int8_t *mat,
                     int32_t o_del, int32_t e_del, int32_t o_ins, int32_t e_ins, int32_t w, int32_t origin,
                     int32_t *Mx, int32_t *Ex, int32_t *Fx) {
    prefix_ctx c = {qlen, tlen, w, query, target, mat, o_del, e_del, o_ins, e_ins, Mx, Ex, Fx};
    for (int32_t k = 0; k < qlen * tlen; ++k) Mx[k] = Ex[k] = Fx[k] = ORC_ENUM_NONE;
    prefix_dfs(&c, -1, -1, origin, 0);
}
