/*
 * SIMD CPU baseline: the reference's inter-sequence form of the two kernels on AVX2, 16 jobs per
 * 256-bit register of int16 lanes.
 *
 * TEST AND BASELINE INFRASTRUCTURE ONLY (see ksw_oracle.c): bench.py times it beside the GPU as the
 * CPU path, and the tests check it against the scalar oracle.  It follows the reference's batched
 * CPU kernels, crates/bwa-neon/src/batched.rs:1370-1841 (define_sw_kernel!: SoA columns
 * [col*LANES + lane], the union band of the lanes with a per-lane mask, the scalar per-lane row
 * epilogue :1763-1823) for the extension and crates/bwa-neon/src/matesw.rs:752-877
 * (fwd_local_sw_batch, with its extract_group :1200-1232) for the rescue's forward and reverse
 * passes, in the lockstep order of the scalar twins restated in ksw_lockstep.c.  Lanes are the
 * jobs of a chunk of 16; a chunk's jobs are neighbours in length order, as the reference bins them.
 * Jobs outside the 16-bit range or with a matrix that is not bwa's uniform DNA one run the scalar
 * functions, as the reference sends them to its scalar kernels (batched.rs:730-734, :836-865).
 */
#include "ksw_oracle.h"

#include <immintrin.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

enum { SL = 16 };          /* lanes: int16 in a 256-bit register */
enum { S_PAD = 255, S_ZPAD = 5 };
static const int32_t S_I16_LIMIT = 30000; /* every value of a 16-bit lane stays below this */

typedef struct {
    size_t *idx;
    size_t n;
} orc_idx_vec;

static int cmp_desc_key(const void *a, const void *b) {
    const uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? 1 : x > y ? -1 : 0;
}

/* Job indices in descending order of (key << 32 | index): neighbours in length share a chunk. */
static size_t *sort_by_key(size_t lo, size_t hi, const uint32_t *k1, const uint32_t *k2) {
    const size_t n = hi - lo;
    uint64_t *kv = (uint64_t *)malloc(n * sizeof(uint64_t) + 1);
    for (size_t r = 0; r < n; ++r) {
        const uint64_t key = ((uint64_t)(k1[lo + r] & 0xffffu) << 16) | (k2[lo + r] & 0xffffu);
        kv[r] = key << 32 | (uint64_t)(lo + r);
    }
    qsort(kv, n, sizeof(uint64_t), cmp_desc_key);
    size_t *idx = (size_t *)malloc(n * sizeof(size_t) + 1);
    for (size_t r = 0; r < n; ++r) idx[r] = (size_t)(kv[r] & 0xffffffffu);
    free(kv);
    return idx;
}

/* ---------------------------------------------------------------------------------------------
 * Extension (ksw_extend2 semantics), 16 jobs in lockstep.
 * ------------------------------------------------------------------------------------------- */
typedef struct {
    const uint8_t *arena;
    const uint64_t *q_off, *t_off;
    const uint32_t *q_len, *t_len;
    const int32_t *h0;
    int32_t m;
    const int8_t *mat;
    int32_t o_del, e_del, o_ins, e_ins, w, end_bonus, zdrop, minsc, max_sc;
    size_t lo, hi;
    void *out;
    int kind; /* 0 extension, 1 rescue */
} simd_slice;

/* One chunk of nl <= 16 extension jobs (uniform DNA matrix, 16-bit range). */
static void ext_chunk(const simd_slice *s, const size_t *jobs, int nl, orc_ext_result *out) {
    const int32_t oe_del = s->o_del + s->e_del, oe_ins = s->o_ins + s->e_ins;
    const int32_t e_del = s->e_del, e_ins = s->e_ins;
    const int32_t a = s->mat[0], mm = s->mat[1], npen = s->mat[s->m * s->m - 1];
    int32_t max_sc = a;
    int32_t ql[SL], tl[SL], h0[SL], w[SL], beg[SL], end[SL], mx[SL], mi[SL], mjb[SL], mie[SL], gs[SL], mo[SL];
    int done[SL];
    int32_t maxq = 0, maxt = 0;
    const uint8_t *qp[SL], *tp[SL];
    for (int l = 0; l < SL; ++l) {
        done[l] = 1;
        ql[l] = tl[l] = 0;
    }
    for (int l = 0; l < nl; ++l) {
        const size_t k = jobs[l];
        ql[l] = (int32_t)s->q_len[k];
        tl[l] = (int32_t)s->t_len[k];
        h0[l] = s->h0[k];
        qp[l] = s->arena + s->q_off[k];
        tp[l] = s->arena + s->t_off[k];
        w[l] = orc_clamp_band(s->w, ql[l], max_sc, s->end_bonus, s->o_ins, s->e_ins, s->o_del, e_del);
        maxq = ql[l] > maxq ? ql[l] : maxq;
        maxt = tl[l] > maxt ? tl[l] : maxt;
    }
    const int32_t S = maxq + 1;
    /* SoA state and query codes: [column][lane] */
    int16_t *EH = (int16_t *)aligned_alloc(32, (size_t)(S + 1) * SL * 2);
    int16_t *EE = (int16_t *)aligned_alloc(32, (size_t)(S + 1) * SL * 2);
    int16_t *QC = (int16_t *)aligned_alloc(32, (size_t)(S + 1) * SL * 2);
    memset(EH, 0, (size_t)(S + 1) * SL * 2);
    memset(EE, 0, (size_t)(S + 1) * SL * 2);
    for (int32_t j = 0; j < S; ++j)
        for (int l = 0; l < SL; ++l) QC[j * SL + l] = (l < nl && j < ql[l]) ? qp[l][j] : 0;
    for (int l = 0; l < nl; ++l) {
        /* row -1 (sw.rs:263-279) */
        EH[l] = (int16_t)h0[l];
        if (ql[l] >= 1) EH[SL + l] = (int16_t)(h0[l] > oe_ins ? h0[l] - oe_ins : 0);
        for (int32_t j = 2; j <= ql[l] && EH[(j - 1) * SL + l] > e_ins; ++j)
            EH[j * SL + l] = (int16_t)(EH[(j - 1) * SL + l] - e_ins);
        mx[l] = h0[l];
        mi[l] = mjb[l] = mie[l] = gs[l] = -1;
        mo[l] = 0;
        beg[l] = 0;
        end[l] = ql[l];
        done[l] = 0;
    }
    const __m256i va = _mm256_set1_epi16((short)a), vmm = _mm256_set1_epi16((short)mm);
    const __m256i vnp = _mm256_set1_epi16((short)npen), vfour = _mm256_set1_epi16(4);
    const __m256i voed = _mm256_set1_epi16((short)oe_del), voei = _mm256_set1_epi16((short)oe_ins);
    const __m256i ved = _mm256_set1_epi16((short)e_del), vei = _mm256_set1_epi16((short)e_ins);
    const __m256i zero = _mm256_setzero_si256();
    int16_t lane_beg[SL], lane_end[SL], lane_t[SL], lane_h1[SL];
    int16_t hrm[SL], hrj[SL], hh1[SL];
    for (int32_t i = 0; i < maxt; ++i) {
        int32_t gb = 1 << 30, ge = 0;
        int any = 0;
        for (int l = 0; l < SL; ++l) {
            const int act = l < nl && !done[l] && i < tl[l];
            lane_t[l] = 0;
            lane_h1[l] = 0;
            if (!act) {
                lane_beg[l] = 0x7fff;  /* no column is in this lane's window */
                lane_end[l] = -1;
                continue;
            }
            any = 1;
            if (beg[l] < i - w[l]) beg[l] = i - w[l];
            if (end[l] > i + w[l] + 1) end[l] = i + w[l] + 1;
            if (end[l] > ql[l]) end[l] = ql[l];
            if (beg[l] == 0) {
                const int32_t v = h0[l] - (s->o_del + e_del * (i + 1));
                lane_h1[l] = (int16_t)(v > 0 ? v : 0);
            }
            lane_beg[l] = (int16_t)beg[l];
            lane_end[l] = (int16_t)end[l];
            lane_t[l] = tp[l][i];
            gb = beg[l] < gb ? beg[l] : gb;
            ge = end[l] > ge ? end[l] : ge;
        }
        if (!any) break;
        const __m256i vbeg = _mm256_loadu_si256((const __m256i *)lane_beg);
        const __m256i vend = _mm256_loadu_si256((const __m256i *)lane_end);
        const __m256i vt = _mm256_loadu_si256((const __m256i *)lane_t);
        const __m256i tn = _mm256_cmpeq_epi16(vt, vfour);
        __m256i h1 = _mm256_loadu_si256((const __m256i *)lane_h1);
        __m256i f = zero, rm = zero, rj = _mm256_set1_epi16(-1);
        for (int32_t j = gb; j < ge; ++j) {
            const __m256i vj = _mm256_set1_epi16((short)j);
            /* in the lane's window: beg <= j < end */
            const __m256i in = _mm256_andnot_si256(_mm256_cmpgt_epi16(vbeg, vj), _mm256_cmpgt_epi16(vend, vj));
            __m256i *ph = (__m256i *)(EH + (size_t)j * SL), *pe = (__m256i *)(EE + (size_t)j * SL);
            __m256i bm = _mm256_load_si256(ph);
            const __m256i e = _mm256_load_si256(pe);
            _mm256_store_si256(ph, _mm256_blendv_epi8(bm, h1, in));
            const __m256i q = _mm256_load_si256((const __m256i *)(QC + (size_t)j * SL));
            __m256i sc = _mm256_blendv_epi8(vmm, va, _mm256_cmpeq_epi16(q, vt));
            sc = _mm256_blendv_epi8(sc, vnp, _mm256_or_si256(tn, _mm256_cmpeq_epi16(q, vfour)));
            /* M = H ? H + s : 0 */
            bm = _mm256_andnot_si256(_mm256_cmpeq_epi16(bm, zero), _mm256_add_epi16(bm, sc));
            const __m256i h = _mm256_max_epi16(_mm256_max_epi16(bm, e), f);
            h1 = _mm256_blendv_epi8(h1, h, in);
            /* row maximum, the later column on ties */
            const __m256i up = _mm256_andnot_si256(_mm256_cmpgt_epi16(rm, h), in);
            rm = _mm256_blendv_epi8(rm, h, up);
            rj = _mm256_blendv_epi8(rj, vj, up);
            const __m256i t1 = _mm256_max_epi16(_mm256_sub_epi16(bm, voed), zero);
            _mm256_store_si256(pe, _mm256_blendv_epi8(e, _mm256_max_epi16(_mm256_sub_epi16(e, ved), t1), in));
            const __m256i t2 = _mm256_max_epi16(_mm256_sub_epi16(bm, voei), zero);
            f = _mm256_blendv_epi8(f, _mm256_max_epi16(_mm256_sub_epi16(f, vei), t2), in);
        }
        _mm256_storeu_si256((__m256i *)hrm, rm);
        _mm256_storeu_si256((__m256i *)hrj, rj);
        _mm256_storeu_si256((__m256i *)hh1, h1);
        /* the per-lane row epilogue (sw.rs:418-480) */
        for (int l = 0; l < nl; ++l) {
            if (done[l] || i >= tl[l]) continue;
            const int32_t rml = hrm[l], rjl = hrj[l], h1l = hh1[l];
            EH[(size_t)end[l] * SL + l] = (int16_t)h1l;
            EE[(size_t)end[l] * SL + l] = 0;
            if (end[l] == ql[l] && gs[l] <= h1l) {
                mie[l] = i;
                gs[l] = h1l;
            }
            if (rml == 0) {
                done[l] = 1;
                continue;
            }
            if (rml > mx[l]) {
                mx[l] = rml;
                mi[l] = i;
                mjb[l] = rjl;
                const int32_t off = abs(rjl - i);
                if (off > mo[l]) mo[l] = off;
            } else if (s->zdrop > 0) {
                const int32_t di = i - mi[l], dj = rjl - mjb[l];
                const int32_t drop = di > dj ? mx[l] - rml - (di - dj) * e_del : mx[l] - rml - (dj - di) * e_ins;
                if (drop > s->zdrop) {
                    done[l] = 1;
                    continue;
                }
            }
            int32_t b = beg[l];
            while (b < end[l] && EH[(size_t)b * SL + l] == 0 && EE[(size_t)b * SL + l] == 0) ++b;
            beg[l] = b;
            int32_t z = end[l];
            while (z >= beg[l] && EH[(size_t)z * SL + l] == 0 && EE[(size_t)z * SL + l] == 0) --z;
            end[l] = z + 2 < ql[l] ? z + 2 : ql[l];
        }
    }
    for (int l = 0; l < nl; ++l) {
        orc_ext_result r = {mx[l], mjb[l] + 1, mi[l] + 1, mie[l] + 1, gs[l], mo[l]};
        out[jobs[l]] = r;
    }
    free(EH);
    free(EE);
    free(QC);
}

static void *ext_slice_run(void *arg) {
    const simd_slice *s = (const simd_slice *)arg;
    orc_ext_result *out = (orc_ext_result *)s->out;
    const int uniform = orc_is_uniform_dna(s->mat, s->m) && s->m == 5;
    int32_t max_sc = 0;
    for (int32_t k = 0; k < s->m * s->m; ++k) max_sc = s->mat[k] > max_sc ? s->mat[k] : max_sc;
    size_t *idx = sort_by_key(s->lo, s->hi, s->q_len, s->t_len);
    size_t *vec = (size_t *)malloc((s->hi - s->lo) * sizeof(size_t) + 1);
    size_t nv = 0;
    for (size_t r = 0; r < s->hi - s->lo; ++r) {
        const size_t k = idx[r];
        const int64_t bound = (int64_t)s->h0[k] + (int64_t)s->q_len[k] * max_sc;
        if (uniform && s->h0[k] > 0 && s->q_len[k] > 0 && bound < S_I16_LIMIT && s->q_len[k] < 8192)
            vec[nv++] = k;
        else
            out[k] = orc_ksw_extend2((int32_t)s->q_len[k], s->arena + s->q_off[k], (int32_t)s->t_len[k],
                                     s->arena + s->t_off[k], s->m, s->mat, s->o_del, s->e_del, s->o_ins, s->e_ins,
                                     s->w, s->end_bonus, s->zdrop, s->h0[k], NULL);
    }
    for (size_t c = 0; c < nv; c += SL) ext_chunk(s, vec + c, (int)(nv - c < SL ? nv - c : SL), out);
    free(vec);
    free(idx);
    return NULL;
}

/* ---------------------------------------------------------------------------------------------
 * Rescue forward / reverse pass (ksw_local_fwd semantics with bwa's padding), 16 jobs in lockstep.
 * Query codes: 0..4, S_ZPAD for padding columns (score 0), S_PAD past a lane's padded length
 * (dead).  Each lane stops once its row maximum reaches endsc (the reverse pass).
 * ------------------------------------------------------------------------------------------- */
typedef struct {
    const uint8_t *q, *t; /* forward order, or already reversed */
    int32_t qlen, tlen, endsc;
} fwd_job;

static void fwd_chunk(const fwd_job *jb, int nl, int32_t m, const int8_t *mat, int32_t o_del, int32_t e_del,
                      int32_t o_ins, int32_t e_ins, int32_t minsc, int32_t max_sc, int want_tracker,
                      orc_fwd_result *res) {
    const int32_t oe_del = o_del + e_del, oe_ins = o_ins + e_ins;
    const int32_t a = mat[0], mmis = mat[1], npen = mat[m * m - 1];
    int32_t qmax = 0, tmax = 0, pad[SL];
    for (int l = 0; l < nl; ++l) {
        pad[l] = orc_padded_qlen(jb[l].qlen, max_sc);
        qmax = pad[l] > qmax ? pad[l] : qmax;
        tmax = jb[l].tlen > tmax ? jb[l].tlen : tmax;
        orc_fwd_result d = {0, -1, -1, -1, -1};
        res[l] = d;
    }
    if (qmax == 0 || tmax == 0) return;
    int16_t *QC = (int16_t *)aligned_alloc(32, (size_t)qmax * SL * 2);
    int16_t *HP = (int16_t *)aligned_alloc(32, (size_t)(qmax + 1) * SL * 2);
    int16_t *EE = (int16_t *)aligned_alloc(32, (size_t)qmax * SL * 2);
    int32_t *rowmax = want_tracker ? (int32_t *)malloc((size_t)tmax * SL * 4) : NULL;
    for (int32_t j = 0; j < qmax; ++j)
        for (int l = 0; l < SL; ++l) {
            int16_t c = S_PAD;
            if (l < nl) c = j < jb[l].qlen ? jb[l].q[j] : j < pad[l] ? S_ZPAD : S_PAD;
            QC[j * SL + l] = c;
        }
    memset(HP, 0, (size_t)(qmax + 1) * SL * 2);
    memset(EE, 0, (size_t)qmax * SL * 2);
    int32_t gmax[SL], te[SL], qe[SL], limit[SL], frozen[SL];
    for (int l = 0; l < SL; ++l) {
        gmax[l] = 0;
        te[l] = -1;
        qe[l] = 0;
        frozen[l] = l >= nl;
        limit[l] = l < nl ? jb[l].tlen - 1 : -1;
    }
    const __m256i va = _mm256_set1_epi16((short)a), vmm = _mm256_set1_epi16((short)mmis);
    const __m256i vnp = _mm256_set1_epi16((short)npen), vfour = _mm256_set1_epi16(4);
    const __m256i vzpad = _mm256_set1_epi16(S_ZPAD), vdead = _mm256_set1_epi16(-30000);
    const __m256i voed = _mm256_set1_epi16((short)oe_del), voei = _mm256_set1_epi16((short)oe_ins);
    const __m256i ved = _mm256_set1_epi16((short)e_del), vei = _mm256_set1_epi16((short)e_ins);
    const __m256i zero = _mm256_setzero_si256();
    int16_t lane_t[SL], himax[SL], hicol[SL];
    for (int32_t i = 0; i < tmax; ++i) {
        int any = 0;
        for (int l = 0; l < SL; ++l) {
            const int act = l < nl && !frozen[l] && i < jb[l].tlen;
            lane_t[l] = act ? jb[l].t[i] : (int16_t)m;  /* a lane past its rows: dead cells */
            any |= act;
        }
        if (!any) break;
        const __m256i vt = _mm256_loadu_si256((const __m256i *)lane_t);
        const __m256i tdead = _mm256_cmpgt_epi16(vt, _mm256_set1_epi16((short)(m - 1)));
        const __m256i tn = _mm256_cmpeq_epi16(vt, vfour);
        __m256i f = zero, hd = zero, imax = zero, icol = zero;
        for (int32_t j = 0; j < qmax; ++j) {
            const __m256i q = _mm256_load_si256((const __m256i *)(QC + (size_t)j * SL));
            __m256i sc = _mm256_blendv_epi8(vmm, va, _mm256_cmpeq_epi16(q, vt));
            sc = _mm256_blendv_epi8(sc, vnp, _mm256_or_si256(tn, _mm256_cmpeq_epi16(q, vfour)));
            sc = _mm256_blendv_epi8(sc, zero, _mm256_cmpeq_epi16(q, vzpad));
            sc = _mm256_blendv_epi8(sc, vdead, _mm256_or_si256(tdead, _mm256_cmpgt_epi16(q, vzpad)));
            __m256i *pe = (__m256i *)(EE + (size_t)j * SL);
            const __m256i e = _mm256_load_si256(pe);
            __m256i h = _mm256_max_epi16(_mm256_adds_epi16(hd, sc), zero);
            h = _mm256_max_epi16(_mm256_max_epi16(h, e), f);
            const __m256i up = _mm256_cmpgt_epi16(h, imax);  /* the first column on ties */
            imax = _mm256_max_epi16(imax, h);
            icol = _mm256_blendv_epi8(icol, _mm256_set1_epi16((short)j), up);
            /* the next column's diagonal is this column's H of the row before */
            __m256i *ph = (__m256i *)(HP + (size_t)j * SL);
            hd = _mm256_load_si256(ph);
            _mm256_store_si256(ph, h);
            const __m256i hoe = _mm256_sub_epi16(h, voed);
            _mm256_store_si256(pe, _mm256_max_epi16(_mm256_max_epi16(_mm256_sub_epi16(e, ved), hoe), zero));
            const __m256i hoi = _mm256_sub_epi16(h, voei);
            f = _mm256_max_epi16(_mm256_max_epi16(_mm256_sub_epi16(f, vei), hoi), zero);
        }
        _mm256_storeu_si256((__m256i *)himax, imax);
        _mm256_storeu_si256((__m256i *)hicol, icol);
        for (int l = 0; l < nl; ++l) {
            if (i >= jb[l].tlen || frozen[l]) continue;
            if (rowmax) rowmax[(size_t)i * SL + l] = himax[l];
            if (himax[l] > gmax[l]) {
                gmax[l] = himax[l];
                te[l] = i;
                qe[l] = hicol[l];
                if (gmax[l] >= jb[l].endsc) {
                    frozen[l] = 1;
                    limit[l] = i;
                }
            }
        }
    }
    int32_t *bs = want_tracker ? (int32_t *)malloc((size_t)(tmax + 1) * 4) : NULL;
    int32_t *br = want_tracker ? (int32_t *)malloc((size_t)(tmax + 1) * 4) : NULL;
    for (int l = 0; l < nl; ++l) {
        orc_fwd_result r;
        r.score = gmax[l];
        r.te = te[l];
        r.qe = te[l] >= 0 ? qe[l] : -1;
        r.score2 = r.te2 = -1;
        if (want_tracker) {
            orc_tracker b = {bs, br, 0};
            for (int32_t i = 0; i <= limit[l]; ++i) orc_tracker_push_row(&b, i, rowmax[(size_t)i * SL + l], minsc);
            orc_tracker_finish(&b, gmax[l], te[l], max_sc, &r.score2, &r.te2);
        }
        res[l] = r;
    }
    free(bs);
    free(br);
    free(rowmax);
    free(QC);
    free(HP);
    free(EE);
}

static void *resc_slice_run(void *arg) {
    const simd_slice *s = (const simd_slice *)arg;
    orc_ksw_result *out = (orc_ksw_result *)s->out;
    const int uniform = orc_is_uniform_dna(s->mat, s->m) && s->m == 5;
    size_t *idx = sort_by_key(s->lo, s->hi, s->q_len, s->t_len);
    const size_t n = s->hi - s->lo;
    size_t *vec = (size_t *)malloc(n * sizeof(size_t) + 1);
    size_t nv = 0;
    for (size_t r = 0; r < n; ++r) {
        const size_t k = idx[r];
        const int32_t ql = (int32_t)s->q_len[k], tl = (int32_t)s->t_len[k];
        const int64_t bound = (int64_t)orc_padded_qlen(ql, s->max_sc) * (s->max_sc > 0 ? s->max_sc : 1);
        if (ql == 0 || tl == 0) {
            orc_ksw_result d = {0, -1, -1, -1, -1, -1, -1};
            out[k] = d;
        } else if (uniform && bound < S_I16_LIMIT && s->mat[0] == s->max_sc) {
            vec[nv++] = k;
        } else {
            out[k] = orc_batched_align2_job(ql, s->arena + s->q_off[k], tl, s->arena + s->t_off[k], s->m, s->mat,
                                            s->o_del, s->e_del, s->o_ins, s->e_ins, s->minsc, s->max_sc, NULL);
        }
    }
    /* forward pass over chunks of neighbours in length; the jobs that qualify are collected for
     * the reverse pass on their reversed prefixes (matesw.rs:388-489) */
    fwd_job jb[SL];
    orc_fwd_result fr[SL];
    size_t *rev = (size_t *)malloc(nv * sizeof(size_t) + 1);
    size_t nr = 0;
    for (size_t c = 0; c < nv; c += SL) {
        const int nl = (int)(nv - c < SL ? nv - c : SL);
        for (int l = 0; l < nl; ++l) {
            const size_t k = vec[c + l];
            jb[l].q = s->arena + s->q_off[k];
            jb[l].t = s->arena + s->t_off[k];
            jb[l].qlen = (int32_t)s->q_len[k];
            jb[l].tlen = (int32_t)s->t_len[k];
            jb[l].endsc = INT32_MAX;
        }
        fwd_chunk(jb, nl, s->m, s->mat, s->o_del, s->e_del, s->o_ins, s->e_ins, s->minsc, s->max_sc, 1, fr);
        for (int l = 0; l < nl; ++l) {
            const size_t k = vec[c + l];
            orc_ksw_result r = {fr[l].score, -1, fr[l].te >= 0 ? fr[l].qe : -1, -1, fr[l].te, fr[l].score2, fr[l].te2};
            out[k] = r;
            if (r.score >= s->minsc && r.qe >= 0) rev[nr++] = k;
        }
    }
    uint8_t *qr[SL], *tr[SL];
    for (size_t c = 0; c < nr; c += SL) {
        const int nl = (int)(nr - c < SL ? nr - c : SL);
        for (int l = 0; l < nl; ++l) {
            const size_t k = rev[c + l];
            const int32_t ql = (int32_t)s->q_len[k];
            const int32_t rq = out[k].qe + 1 < ql ? out[k].qe + 1 : ql, rt = out[k].te + 1;
            const uint8_t *q = s->arena + s->q_off[k], *t = s->arena + s->t_off[k];
            qr[l] = (uint8_t *)malloc((size_t)rq + 1);
            tr[l] = (uint8_t *)malloc((size_t)rt + 1);
            for (int32_t x = 0; x < rq; ++x) qr[l][x] = q[rq - 1 - x];
            for (int32_t x = 0; x < rt; ++x) tr[l][x] = t[rt - 1 - x];
            jb[l].q = qr[l];
            jb[l].t = tr[l];
            jb[l].qlen = rq;
            jb[l].tlen = rt;
            jb[l].endsc = out[k].score;
        }
        fwd_chunk(jb, nl, s->m, s->mat, s->o_del, s->e_del, s->o_ins, s->e_ins, INT32_MAX, s->max_sc, 0, fr);
        for (int l = 0; l < nl; ++l) {
            const size_t k = rev[c + l];
            if (fr[l].score == out[k].score) {
                out[k].tb = out[k].te - fr[l].te;
                out[k].qb = out[k].qe - fr[l].qe;
            }
            free(qr[l]);
            free(tr[l]);
        }
    }
    free(rev);
    free(vec);
    free(idx);
    return NULL;
}

static void run_simd_slices(simd_slice *proto, size_t n, int nthreads, void *(*fn)(void *)) {
    if (nthreads < 1) nthreads = 1;
    if ((size_t)nthreads > n && n > 0) nthreads = (int)n;
    simd_slice *sl = (simd_slice *)calloc((size_t)nthreads, sizeof(simd_slice));
    pthread_t *th = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
    for (int t = 0; t < nthreads; ++t) {
        sl[t] = *proto;
        sl[t].lo = n * (size_t)t / (size_t)nthreads;
        sl[t].hi = n * (size_t)(t + 1) / (size_t)nthreads;
        if (nthreads == 1) fn(&sl[t]);
        else pthread_create(&th[t], NULL, fn, &sl[t]);
    }
    for (int t = 0; t < nthreads && nthreads > 1; ++t) pthread_join(th[t], NULL);
    free(sl);
    free(th);
}

void orc_simd_extend_batch(size_t n, const uint8_t *arena, const uint64_t *q_off, const uint32_t *q_len,
                           const uint64_t *t_off, const uint32_t *t_len, const int32_t *h0, int32_t m,
                           const int8_t *mat, int32_t o_del, int32_t e_del, int32_t o_ins, int32_t e_ins,
                           int32_t w, int32_t end_bonus, int32_t zdrop, orc_ext_result *out, int nthreads) {
    simd_slice p;
    memset(&p, 0, sizeof p);
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
    run_simd_slices(&p, n, nthreads, ext_slice_run);
}

void orc_simd_align2_batch(size_t n, const uint8_t *arena, const uint64_t *q_off, const uint32_t *q_len,
                           const uint64_t *t_off, const uint32_t *t_len, int32_t m, const int8_t *mat,
                           int32_t o_del, int32_t e_del, int32_t o_ins, int32_t e_ins, int32_t minsc,
                           int32_t max_sc, orc_ksw_result *out, int nthreads) {
    simd_slice p;
    memset(&p, 0, sizeof p);
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
    p.out = out;
    run_simd_slices(&p, n, nthreads, resc_slice_run);
}
