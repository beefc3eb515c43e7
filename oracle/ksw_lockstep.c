/*
 * Second, independent CPU restatement of the two kernels in the reference's lockstep form.
 *
 * TEST INFRASTRUCTURE ONLY (see ksw_oracle.c).  Where ksw_oracle.c follows the per-job scalar code,
 * this file follows the reference's inter-sequence lockstep twins, which process `lanes` jobs with
 * shared row/column loops over SoA arrays indexed [position*lanes + lane]:
 *
 *   orc_lockstep_extend     crates/bwa-neon/src/batched.rs:947-1251 (batched_extend_scalar)
 *   orc_lockstep_local_fwd  crates/bwa-neon/src/matesw.rs:933-1161 (fwd_local_sw_scalar) and
 *                           :1200-1232 (extract_group)
 *
 * The two forms differ in structure (union band with per-lane masks, PAD/ZPAD sequence codes and a
 * dead-cell score instead of explicit padding logic, a per-row max array replayed into the tracker
 * after the DP, the argmax column tracked online instead of a row snapshot), so agreement between
 * them checks each against the other.
 */
#include "ksw_oracle.h"

#include <stdlib.h>
#include <string.h>

enum { LS_PAD = 255, LS_ZPAD = 5 };
static const int32_t LS_DEAD = -30000; /* matesw.rs:634 */

static int32_t ls_max_sc(int32_t m, const int8_t *mat) {
    int32_t mx = mat[0];
    for (int32_t k = 1; k < m * m; ++k) mx = mat[k] > mx ? mat[k] : mx;
    return mx;
}

void orc_lockstep_extend(size_t n, const uint8_t *const *queries, const int32_t *qlens,
                         const uint8_t *const *targets, const int32_t *tlens, const int32_t *h0s,
                         int32_t m, const int8_t *mat, int32_t o_del, int32_t e_del,
                         int32_t o_ins, int32_t e_ins, int32_t w0, int32_t end_bonus,
                         int32_t zdrop, int32_t L, orc_ext_result *out) {
    const int32_t oe_del = o_del + e_del, oe_ins = o_ins + e_ins;
    const int32_t max_sc = ls_max_sc(m, mat);
    int32_t *ql = calloc(L, 4), *tl = calloc(L, 4), *h0 = calloc(L, 4), *w = calloc(L, 4);
    int32_t *beg = calloc(L, 4), *end = calloc(L, 4), *mx = calloc(L, 4), *mi = calloc(L, 4);
    int32_t *mjb = calloc(L, 4), *mie = calloc(L, 4), *gs = calloc(L, 4), *mo = calloc(L, 4);
    int32_t *h1 = calloc(L, 4), *f = calloc(L, 4), *rm = calloc(L, 4), *rj = calloc(L, 4);
    char *done = calloc(L, 1), *act = calloc(L, 1);
    for (size_t c0 = 0; c0 < n; c0 += (size_t)L) {
        int32_t nl = (int32_t)((n - c0) < (size_t)L ? (n - c0) : (size_t)L);
        int32_t maxq = 0, maxt = 0;
        for (int32_t l = 0; l < nl; ++l) {
            ql[l] = qlens[c0 + l];
            tl[l] = tlens[c0 + l];
            h0[l] = h0s[c0 + l];
            w[l] = orc_clamp_band(w0, ql[l], max_sc, end_bonus, o_ins, e_ins, o_del, e_del);
            maxq = ql[l] > maxq ? ql[l] : maxq;
            maxt = tl[l] > maxt ? tl[l] : maxt;
        }
        const int32_t S = maxq + 1;
        int32_t *EH = calloc((size_t)L * S, 4), *EE = calloc((size_t)L * S, 4);
        for (int32_t l = 0; l < L; ++l) done[l] = 1;
        for (int32_t l = 0; l < nl; ++l) {
            EH[l * S] = h0[l];
            if (ql[l] >= 1) EH[l * S + 1] = h0[l] > oe_ins ? h0[l] - oe_ins : 0;
            for (int32_t j = 2; j <= ql[l] && EH[l * S + j - 1] > e_ins; ++j)
                EH[l * S + j] = EH[l * S + j - 1] - e_ins;
            mx[l] = h0[l];
            mi[l] = mjb[l] = mie[l] = gs[l] = -1;
            mo[l] = 0;
            beg[l] = 0;
            end[l] = ql[l];
            done[l] = 0;
        }
        for (int32_t i = 0; i < maxt; ++i) {
            int32_t gb = 1 << 30, ge = 0;
            for (int32_t l = 0; l < nl; ++l) {
                act[l] = !done[l] && i < tl[l];
                if (!act[l]) continue;
                f[l] = 0;
                rm[l] = 0;
                rj[l] = -1;
                if (beg[l] < i - w[l]) beg[l] = i - w[l];
                if (end[l] > i + w[l] + 1) end[l] = i + w[l] + 1;
                if (end[l] > ql[l]) end[l] = ql[l];
                h1[l] = 0;
                if (beg[l] == 0) {
                    int32_t v = h0[l] - (o_del + e_del * (i + 1));
                    h1[l] = v > 0 ? v : 0;
                }
                gb = beg[l] < gb ? beg[l] : gb;
                ge = end[l] > ge ? end[l] : ge;
            }
            for (int32_t j = gb; j < ge; ++j) {
                for (int32_t l = 0; l < nl; ++l) {
                    if (!act[l] || j < beg[l] || j >= end[l]) continue;
                    int32_t *ph = &EH[l * S + j], *pe = &EE[l * S + j];
                    int32_t bm = *ph, e = *pe;
                    *ph = h1[l];
                    int32_t sc = mat[targets[c0 + l][i] * m + queries[c0 + l][j]];
                    bm = bm != 0 ? bm + sc : 0;
                    int32_t h = bm > e ? bm : e;
                    h = h > f[l] ? h : f[l];
                    h1[l] = h;
                    if (rm[l] <= h) { rj[l] = j; rm[l] = h; }
                    int32_t t = bm - oe_del > 0 ? bm - oe_del : 0;
                    *pe = e - e_del > t ? e - e_del : t;
                    t = bm - oe_ins > 0 ? bm - oe_ins : 0;
                    f[l] = f[l] - e_ins > t ? f[l] - e_ins : t;
                }
            }
            for (int32_t l = 0; l < nl; ++l) {
                if (!act[l]) continue;
                int32_t *H = EH + l * S, *E = EE + l * S;
                H[end[l]] = h1[l];
                E[end[l]] = 0;
                if (end[l] == ql[l] && gs[l] <= h1[l]) { mie[l] = i; gs[l] = h1[l]; }
                if (rm[l] == 0) { done[l] = 1; continue; }
                if (rm[l] > mx[l]) {
                    mx[l] = rm[l];
                    mi[l] = i;
                    mjb[l] = rj[l];
                    int32_t off = abs(rj[l] - i);
                    if (off > mo[l]) mo[l] = off;
                } else if (zdrop > 0) {
                    int32_t drop = (i - mi[l] > rj[l] - mjb[l])
                                       ? mx[l] - rm[l] - ((i - mi[l]) - (rj[l] - mjb[l])) * e_del
                                       : mx[l] - rm[l] - ((rj[l] - mjb[l]) - (i - mi[l])) * e_ins;
                    if (drop > zdrop) { done[l] = 1; continue; }
                }
                int32_t a = beg[l];
                while (a < end[l] && H[a] == 0 && E[a] == 0) ++a;
                beg[l] = a;
                int32_t z = end[l];
                while (z >= beg[l] && H[z] == 0 && E[z] == 0) --z;
                end[l] = z + 2 < ql[l] ? z + 2 : ql[l];
            }
        }
        for (int32_t l = 0; l < nl; ++l) {
            orc_ext_result r = {mx[l], mjb[l] + 1, mi[l] + 1, mie[l] + 1, gs[l], mo[l]};
            out[c0 + l] = r;
        }
        free(EH);
        free(EE);
    }
    free(ql); free(tl); free(h0); free(w); free(beg); free(end); free(mx); free(mi);
    free(mjb); free(mie); free(gs); free(mo); free(h1); free(f); free(rm); free(rj);
    free(done); free(act);
}

void orc_lockstep_local_fwd(size_t n, const uint8_t *const *queries, const int32_t *qlens,
                            const uint8_t *const *targets, const int32_t *tlens,
                            const int32_t *minsc, const int32_t *endsc, int32_t m,
                            const int8_t *mat, int32_t o_del, int32_t e_del, int32_t o_ins,
                            int32_t e_ins, int32_t max_sc, int32_t L, orc_fwd_result *out) {
    const int32_t oe_del = o_del + e_del, oe_ins = o_ins + e_ins;
    for (size_t c0 = 0; c0 < n; c0 += (size_t)L) {
        int32_t nl = (int32_t)((n - c0) < (size_t)L ? (n - c0) : (size_t)L);
        int32_t qmax = 0, tmax = 0;
        for (int32_t l = 0; l < nl; ++l) {
            int32_t p = orc_padded_qlen(qlens[c0 + l], max_sc);
            qmax = p > qmax ? p : qmax;
            tmax = tlens[c0 + l] > tmax ? tlens[c0 + l] : tmax;
        }
        for (int32_t l = 0; l < nl; ++l) {
            orc_fwd_result d = {0, -1, -1, -1, -1};
            out[c0 + l] = d;
        }
        if (qmax == 0 || tmax == 0) continue;
        uint8_t *sq = malloc((size_t)qmax * L), *st = malloc((size_t)tmax * L);
        memset(sq, LS_PAD, (size_t)qmax * L);
        memset(st, LS_PAD, (size_t)tmax * L);
        for (int32_t l = 0; l < nl; ++l) {
            int32_t q = qlens[c0 + l], p = orc_padded_qlen(q, max_sc);
            for (int32_t c = 0; c < q; ++c) sq[c * L + l] = queries[c0 + l][c];
            for (int32_t c = q; c < p; ++c) sq[c * L + l] = LS_ZPAD;
            for (int32_t r = 0; r < tlens[c0 + l]; ++r) st[r * L + l] = targets[c0 + l][r];
        }
        int32_t *hp = calloc((size_t)qmax * L, 4), *hc = calloc((size_t)qmax * L, 4);
        int32_t *ee = calloc((size_t)qmax * L, 4), *rowmax = calloc((size_t)tmax * L, 4);
        int32_t *gmax = calloc(L, 4), *te = malloc(L * 4), *qe = calloc(L, 4);
        int32_t *limit = malloc(L * 4), *f = malloc(L * 4), *hd = malloc(L * 4);
        int32_t *imax = malloc(L * 4), *icol = malloc(L * 4);
        char *frozen = calloc(L, 1);
        for (int32_t l = 0; l < L; ++l) { te[l] = -1; limit[l] = l < nl ? tlens[c0 + l] - 1 : -1; }
        for (int32_t i = 0; i < tmax; ++i) {
            for (int32_t l = 0; l < L; ++l) { f[l] = 0; hd[l] = 0; imax[l] = 0; icol[l] = 0; }
            for (int32_t j = 0; j < qmax; ++j) {
                for (int32_t l = 0; l < L; ++l) {
                    int32_t tc = st[i * L + l], qc = sq[j * L + l];
                    int32_t sc = (tc >= m || qc > LS_ZPAD) ? LS_DEAD
                                 : (qc == LS_ZPAD)         ? 0
                                                           : mat[tc * m + qc];
                    int32_t cell = j * L + l;
                    int32_t h = hd[l] + sc;
                    if (h < 0) h = 0;
                    if (ee[cell] > h) h = ee[cell];
                    if (f[l] > h) h = f[l];
                    if (h > imax[l]) { imax[l] = h; icol[l] = j; }
                    hc[cell] = h;
                    int32_t en = ee[cell] - e_del;
                    if (h - oe_del > en) en = h - oe_del;
                    ee[cell] = en > 0 ? en : 0;
                    int32_t fn = f[l] - e_ins;
                    if (h - oe_ins > fn) fn = h - oe_ins;
                    f[l] = fn > 0 ? fn : 0;
                    hd[l] = hp[cell];
                }
            }
            for (int32_t l = 0; l < nl; ++l) {
                if (i >= tlens[c0 + l] || frozen[l]) continue;
                rowmax[i * L + l] = imax[l];
                if (imax[l] > gmax[l]) {
                    gmax[l] = imax[l];
                    te[l] = i;
                    qe[l] = icol[l];
                    if (gmax[l] >= endsc[c0 + l]) { frozen[l] = 1; limit[l] = i; }
                }
            }
            int32_t *tmp = hp; hp = hc; hc = tmp;
        }
        /* extract_group (matesw.rs:1200-1232): replay row maxima into the tracker */
        int32_t *bs = malloc((size_t)(tmax + 1) * 4), *br = malloc((size_t)(tmax + 1) * 4);
        for (int32_t l = 0; l < nl; ++l) {
            orc_tracker b = {bs, br, 0};
            for (int32_t i = 0; i <= limit[l]; ++i)
                orc_tracker_push_row(&b, i, rowmax[i * L + l], minsc[c0 + l]);
            orc_fwd_result r;
            r.score = gmax[l];
            r.te = te[l];
            r.qe = te[l] >= 0 ? qe[l] : -1;
            orc_tracker_finish(&b, gmax[l], te[l], max_sc, &r.score2, &r.te2);
            out[c0 + l] = r;
        }
        free(bs); free(br);
        free(sq); free(st); free(hp); free(hc); free(ee); free(rowmax); free(gmax); free(te);
        free(qe); free(limit); free(f); free(hd); free(imax); free(icol); free(frozen);
    }
}
