/*
 * Oracle restatement of the reference's ungapped-diagonal fast path for seed extension (test
 * infrastructure only).  bwa-mem4's batched extension entry point first tries a closed-form walk
 * along the main diagonal and runs the banded DP only for the jobs it declines:
 *
 *   orc_is_uniform_dna    crates/bwa-neon/src/batched.rs:493-521 (is_uniform_dna)
 *   orc_ungapped_xmax     crates/bwa-neon/src/batched.rs:756-764 (x_threshold in simd_dispatch)
 *   orc_ungapped_hit      crates/bwa-neon/src/batched.rs:602-697 (ungapped_hit)
 *
 * The reference states the accepted results are identical to ksw_extend2's; the oracle tests
 * check that claim against orc_ksw_extend2 on every accepted job of the workloads and gate sweeps.
 */
#include <stdint.h>
#include <stddef.h>

#include "ksw_oracle.h"

int orc_is_uniform_dna(const int8_t *mat, int32_t m) {
    if (m < 2) return 0;
    const int8_t a = mat[0], mm = mat[1], npen = mat[m - 1];
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) {
            const int8_t want = (i == m - 1 || j == m - 1) ? npen : (i == j ? a : mm);
            if (mat[i * m + j] != want) return 0;
        }
    return 1;
}

/* Largest number of mismatches for which the ungapped diagonal is provably the strict optimum:
 * X * (a + b) < o_min + e_min, with a the match score and b the mismatch magnitude; -1 when the
 * matrix is degenerate (a + b <= 0). */
int32_t orc_ungapped_xmax(const int8_t *mat, int32_t o_del, int32_t e_del, int32_t o_ins, int32_t e_ins) {
    const int32_t a = mat[0], b = -(int32_t)mat[1];
    const int32_t o_min = o_del < o_ins ? o_del : o_ins, e_min = e_del < e_ins ? e_del : e_ins;
    const int32_t denom = a + b;
    return denom > 0 ? (o_min + e_min - 1) / denom : -1;
}

/* Returns 1 and fills *out when the closed form applies, 0 when the job needs the banded DP. */
int orc_ungapped_hit(int32_t qlen, const uint8_t *query, int32_t tlen, const uint8_t *target, int32_t h0,
                     int32_t a, int32_t b, int32_t xmax, int32_t zdrop, orc_ext_result *out) {
    if (qlen <= 0 || h0 <= 0 || tlen < qlen || xmax < 0) return 0;
    int32_t nmis = 0;
    int32_t pos[128];  /* mismatch columns, all below 128 when the walk is accepted */
    for (int32_t c = 0; c < qlen; ++c) {
        const uint8_t qb = query[c], tb = target[c];
        if (qb >= 4 || tb >= 4) return 0;
        if (qb != tb) {
            if (++nmis > xmax) return 0;
            if (c >= 128) return 0;
            pos[nmis - 1] = c;
        }
    }
    if (nmis == 0) {
        const int32_t s = h0 + qlen * a;
        out->score = s;
        out->qle = out->tle = out->gtle = qlen;
        out->gscore = s;
        out->max_off = 0;
        return 1;
    }
    if (qlen > 128) return 0;
    /* the reference's per-column walk: match +a, mismatch -b floored at 0, strict best update,
     * z-drop declines, a dead diagonal stays dead */
    int32_t running = h0, best = h0, best_end = 0, k = 0;
    for (int32_t c = 0; c < qlen; ++c) {
        if (running == 0) continue;
        if (k < nmis && pos[k] == c) {
            ++k;
            running -= b;
            if (running < 0) running = 0;
        } else {
            running += a;
        }
        if (running > best) {
            best = running;
            best_end = c + 1;
        } else if (zdrop > 0 && best - running > zdrop) {
            return 0;
        }
    }
    if (running == 0) return 0;
    out->score = best;
    out->qle = out->tle = best_end;
    out->gtle = qlen;
    out->gscore = running;
    out->max_off = 0;
    return 1;
}

void orc_ungapped_batch(size_t n, const uint8_t *arena, const uint64_t *q_off, const uint32_t *q_len,
                        const uint64_t *t_off, const uint32_t *t_len, const int32_t *h0, int32_t m,
                        const int8_t *mat, int32_t o_del, int32_t e_del, int32_t o_ins, int32_t e_ins,
                        int32_t zdrop, orc_ext_result *out, uint8_t *hit) {
    const int uniform = orc_is_uniform_dna(mat, m);
    const int32_t xmax = orc_ungapped_xmax(mat, o_del, e_del, o_ins, e_ins);
    for (size_t k = 0; k < n; ++k) {
        hit[k] = uniform ? (uint8_t)orc_ungapped_hit((int32_t)q_len[k], arena + q_off[k], (int32_t)t_len[k],
                                                     arena + t_off[k], h0[k], mat[0], -(int32_t)mat[1], xmax,
                                                     zdrop, &out[k])
                         : 0;
    }
}
