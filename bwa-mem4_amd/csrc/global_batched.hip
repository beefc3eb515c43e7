// Banded global alignment with traceback on the device: ksw_global2 semantics
// (crates/bwa-extend/src/sw.rs:561-740), the DP that turns a settled region into its CIGAR.
//
// Two kernels per batch:
//
//  * glb_group_kernel<K>: one job per 16-lane DPP row, the same mapping as the extension's group
//    kernel.  Row i's band [max(0, i-w), min(i+w+1, qlen)) is fixed by w; lane g owns the K window
//    columns beg + g*K .. beg + g*K + K-1.  H (shifted by one column, as the reference's eh_h) and
//    E live in LDS over absolute columns; F is the max-plus recurrence F(j+1) = max(F(j) - e_ins,
//    M(j) - oe_ins) with no floor, evaluated by a per-lane pass, a max-scan across the row and a
//    second pass.  Every evaluated cell leaves a 4-bit traceback code in HBM: bit 0 "E beat M",
//    bit 1 "F beat max(M, E)" (so H came from M, E or F with ties M over E over F), bit 2 "the
//    deletion continues", bit 3 "the insertion continues", the same information as the
//    reference's byte f<<4 | e<<2 | h.  Each bit is the sign of a difference, shifted into the
//    lane's code word with one funnel shift.  Lane g's K codes of a row fill (K+7)/8 dwords; in
//    the dword holding n codes, code kk (0-based) is the nibble n-1-kk.  Rows are stored in tiles
//    of 8: dword m of lane g in row i sits at ((i/8*16 + g) * dwords + m) * 8 + i%8, so the 8 rows
//    of one lane's dword are adjacent and a traceback step along a diagonal (same lane, previous
//    row) mostly reads the cache line of the step before.
//  * glb_traceback_kernel: one job per thread walks the codes back from the band's last cell,
//    merging runs as push_cigar does (sw.rs:515-524), and leaves the runs back to front in a
//    per-job scratch region; glb_cigar_gather_kernel copies them front to back into the caller's
//    compact layout.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>

#include "bwg_batched_kernels.h"
#include "dpp.h"

namespace bwg {
namespace {

constexpr int32_t kMinusInf = -0x40000000;  // the reference's MINUS_INF (sw.rs:499)
constexpr int32_t kNeg = -0x60000000;       // below every reachable value: the F chain's identity
constexpr int kGlbGroupsPerBlock = 16;
constexpr uint32_t kTracebackOffBand = 0xffffffffu;

__host__ __device__ __forceinline__ int glb_eh_bytes(int qcap) { return 8 * (qcap + 2); }
// Scores per query column: five int8 planes (mat[c][q[j]] for target base c), or with packed
// scores one byte 6*q[j] that selects a 6-bit field of the row's packed score word.
__host__ __device__ __forceinline__ int glb_plane_bytes(int qcap, bool packed) { return packed ? qcap : 5 * qcap; }
// Block LDS: the packed score words (32 bytes, packed form only), then every group's H/E columns,
// then every group's score bytes, then slack for the last group's loads past its columns.
__host__ __device__ __forceinline__ size_t glb_block_lds_bytes(int qcap, int K, bool packed) {
    return (size_t)kGlbGroupsPerBlock * (glb_eh_bytes(qcap) + glb_plane_bytes(qcap, packed)) + 16 * K * 8 + 64 +
           (packed ? 32 : 0);
}

// Position of dword m of lane g's codes in row i, in a job's code region (rows in tiles of 8).
__device__ __forceinline__ size_t code_index(int32_t i, int32_t g, int32_t m, int32_t dw) {
    return ((size_t)((i >> 3) * 16 + g) * dw + m) * 8 + (i & 7);
}

// Shift the sign bit of d into the low end of w: ({w, d} >> 31), one v_alignbit_b32.
__device__ __forceinline__ uint32_t push_sign(uint32_t w, int32_t d) {
    return __builtin_amdgcn_alignbit(w, (uint32_t)d, 31u);
}

// Inclusive prefix max over the 16 lanes of a row for values of any sign: a lane whose source is
// outside the row keeps its own value.
__device__ __forceinline__ int row_scan_max_any(int x) {
    x = max(x, dpp<kRowShr + 1, false>(x, x));
    x = max(x, dpp<kRowShr + 2, false>(x, x));
    x = max(x, dpp<kRowShr + 4, false>(x, x));
    x = max(x, dpp<kRowShr + 8, false>(x, x));
    return x;
}

template <int K, bool PACKED>
__global__ __launch_bounds__(256) void glb_group_kernel(const uint8_t *__restrict__ arena,
                                                        const uint8_t *__restrict__ ref,
                                                        const GlobalJob *__restrict__ jobs,
                                                        const uint32_t *__restrict__ order, uint32_t count,
                                                        uint32_t *__restrict__ head, GlobalScoring sc, int qcap,
                                                        uint32_t *__restrict__ z, int32_t *__restrict__ score) {
    extern __shared__ int32_t lds[];
    constexpr int DW = (K + 7) / 8;
    const int lane = threadIdx.x & 63;
    const int g = lane & 15;
    const int gid = threadIdx.x >> 4;
    char *lb = (char *)lds + (PACKED ? 32 : 0);
    uint32_t *tabs = (uint32_t *)lds;  // packed form: row c's scores mat[c][0..4] as 6-bit fields
    int32_t *eh = (int32_t *)(lb + gid * glb_eh_bytes(qcap));
    int8_t *planes = (int8_t *)lb + kGlbGroupsPerBlock * glb_eh_bytes(qcap) + gid * glb_plane_bytes(qcap, PACKED);
    if constexpr (PACKED) {
        if (threadIdx.x < 5) {
            uint32_t tw = 0;
            for (int c = 0; c < 5; ++c) tw |= ((uint32_t)sc.mat[threadIdx.x * 5 + c] & 63u) << (6 * c);
            tabs[threadIdx.x] = tw;
        }
        __syncthreads();
    }
    const int32_t oe_del = sc.o_del + sc.e_del, oe_ins = sc.o_ins + sc.e_ins;
    const int32_t e_del = sc.e_del, e_ins = sc.e_ins;
    const int32_t Ke = K * e_ins;

    bool active = false, exhausted = false;
    TargetView t = {nullptr, 1};
    int32_t qlen = 0, tlen = 0, w = 0, i = 0, tcur = 0;
    uint32_t idx = 0;
    uint32_t *zj = nullptr;
    // the lane's codes of the row before, not yet stored, for its pn window columns (0: none)
    int32_t pn = 0;
    uint32_t *pzr = nullptr;
    uint32_t *zrow = nullptr;  // this row's first code dword of the lane (code_index(i, g, 0, DW))
    int32_t h1d = 0;           // -(o_del + e_del*(i+1)), carried row to row
    uint32_t pzw[DW];
#pragma unroll
    for (int m = 0; m < DW; ++m) pzw[m] = 0u;
    GroupSlots slots(count, true);
    for (;;) {
        if (!active && !exhausted) {
            const uint32_t slot = slots.take(head);
            if (slot >= count) {
                exhausted = true;
            } else {
                idx = order[slot];
                const GlobalJob jb = jobs[idx];
                const uint8_t *q = arena + jb.q_off;
                t = target_view(arena, ref, jb.t_off, jb.tlen);
                qlen = (int32_t)jb.qlen;
                tlen = (int32_t)jb.tlen;
                w = max(jb.w, 0);
                zj = z + jb.z_off;
                // row -1 (sw.rs:614-621): H(-1, j-1) is a leading insertion of length j within the
                // band, unreachable beyond it; no E is open
                const int32_t wq = min(qlen, w);
                for (int j = g; j <= qlen + 1; j += 16) {
                    const int32_t h = j == 0 ? 0 : (j <= wq ? -(sc.o_ins + sc.e_ins * j) : kMinusInf);
                    *(int2 *)(eh + 2 * j) = make_int2(h, kMinusInf);
                }
                for (int j = g; j < qlen; j += 16) {
                    const int qc = q[j];
                    if constexpr (PACKED) {
                        planes[j] = (int8_t)(6 * qc);
                    } else {
#pragma unroll
                        for (int c = 0; c < 5; ++c) planes[c * qcap + j] = sc.mat[c * 5 + qc];
                    }
                }
                i = 0;
                h1d = -(sc.o_del + sc.e_del);  // H(i, -1) while the band touches column 0, row 0
                zrow = zj + code_index(0, g, 0, DW);
                tcur = tlen > 0 ? t[0] : 0;
                active = tlen > 0;
                if (!active && g == 0)
                    score[idx] = qlen == 0 ? 0 : (qlen <= w ? -(sc.o_ins + sc.e_ins * qlen) : kMinusInf);
            }
        }
        wave_lds_sync();
        if (__ballot(active || !exhausted) == 0) break;
        if (!active) continue;

        // ---- one DP row ----------------------------------------------------------------------
        const int32_t beg = max(i - w, 0);
        const int32_t end = min(i + w + 1, qlen);
        const int32_t h1_init = beg == 0 ? h1d : kMinusInf;  // -(o_del + e_del*(i+1)) at column 0
        const int32_t base = beg + g * K;
        const int32_t nval = min(max(end - base, 0), K);
        const int2 *ehb = (const int2 *)eh + base;
        const int8_t *sb = planes + (PACKED ? 0 : tcur * qcap) + base;
        const uint32_t rowtab = PACKED ? tabs[tcur] : 0u;
        // The vector-memory counter is in order and counts stores too: this row's target base
        // (requested a row ago) is consumed first, then the previous row's codes are stored, then
        // the next base is requested, so the wait for a base never covers a request or a store made
        // just before it.
        asm volatile("" ::"v"(tcur) : "memory");
        if (pn > 0) {
#pragma unroll
            for (int m = 0; m < DW; ++m)
                if (8 * m < pn) pzr[8 * m] = pzw[m];
        }
        const int32_t tnext = i + 1 < tlen ? t[i + 1] : 0;
        int2 v[K];
        int32_t s[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            v[k] = ehb[k];
            if constexpr (PACKED)
                s[k] = __builtin_amdgcn_sbfe(rowtab, (uint32_t)(uint8_t)sb[k], 6u);
            else
                s[k] = sb[k];
        }
        // pass 1: M and the lane's own F chain from the identity, carried with its drift
        // G_k = F_k + k*e_ins: F_{k+1} = max(F_k - e_ins, M_k - oe_ins) becomes G_{k+1} =
        // max(G_k, M_k + (k+1)*e_ins - oe_ins), one add and one max per column
        int32_t mm[K], gg[K];
        int32_t grun = kNeg;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            mm[k] = v[k].x + s[k];
            gg[k] = grun;
            grun = max(grun, mm[k] + ((k + 1) * e_ins - oe_ins));
        }
        grun -= Ke;  // F leaving the lane's block
        // F entering this lane: the row starts at F = MINUS_INF, then the max-scan of the lanes'
        // exits shifted by K*e_ins per lane
        const int32_t incl = row_scan_max_any(grun + g * Ke);
        const int32_t excl = row_shift_in(incl, kNeg);
        const int32_t f_in = g == 0 ? kMinusInf : max(kMinusInf - g * Ke, excl - (g - 1) * Ke);
        uint32_t zw[DW];
#pragma unroll
        for (int m = 0; m < DW; ++m) zw[m] = 0u;
        int2 *ehw = (int2 *)eh + base;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int32_t fd = max(f_in, gg[k]);  // F(i, j) + k*e_ins
            const int32_t f = fd - k * e_ins;
            const int32_t e = v[k].y;
            const int32_t hme = max(mm[k], e);
            const int32_t h = max(hme, f);
            const int32_t td = mm[k] - oe_del, en = e - e_del;
            // The code bits as signs of differences; every value here lies within 2^30 of zero,
            // so no difference overflows.  Bit 3: F(i, j+1) extends F(i, j) rather than opening
            // from M, F - e_ins > M - oe_ins with both sides moved by (k+1)*e_ins.  Bit 2: E(i+1, j)
            // extends E(i, j).  Bits 1 and 0: F beat max(M, E), E beat M; M wins ties over E and
            // both over F (sw.rs:654-661).
            uint32_t w = zw[k / 8];
            w = push_sign(w, mm[k] + ((k + 1) * e_ins - oe_ins) - fd);
            w = push_sign(w, td - en);
            w = push_sign(w, hme - f);
            w = push_sign(w, mm[k] - e);
            zw[k / 8] = w;
            // expected taken: the store runs under the exec mask with no branch around it
            if (__builtin_expect(k < nval, 1)) {
                ehw[k].y = max(en, td);  // E(i+1, j)
                ehw[k + 1].x = h;        // H(i, j), the diagonal of column j+1 next row
            }
        }
        if (g == 0) {
            eh[2 * beg] = h1_init;        // H(i, beg-1)
            eh[2 * end + 1] = kMinusInf;  // E at the band's end
        }
        // this row's codes are stored at the start of the next row, or now if the job ends
        pn = nval;
        pzr = zrow;
        // next row: the next dword of the tile, or the first row of the next tile
        zrow += ((i + 1) & 7) ? 1 : 16 * DW * 8 - 7;
        h1d -= e_del;
#pragma unroll
        for (int m = 0; m < DW; ++m) pzw[m] = zw[m];
        wave_lds_sync();
        ++i;
        tcur = tnext;
        if (i >= tlen) {
            if (pn > 0) {
#pragma unroll
                for (int m = 0; m < DW; ++m)
                    if (8 * m < pn) pzr[8 * m] = pzw[m];
            }
            pn = 0;
            if (g == 0) score[idx] = eh[2 * qlen];  // H(tlen-1, qlen-1) as the last row left it
            active = false;
        }
    }
}

// Walk the traceback codes of one job (sw.rs:650-688); runs are left back to front.
__global__ __launch_bounds__(256) void glb_traceback_kernel(const GlobalJob *__restrict__ jobs, uint32_t n,
                                                            const uint32_t *__restrict__ z,
                                                            uint32_t *__restrict__ ops,
                                                            uint32_t *__restrict__ n_ops) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const GlobalJob J = jobs[k];
    if (J.k == 0) return;  // a long query: glb_plain_kernel walked its own codes
    const int32_t w = max(J.w, 0), qlen = (int32_t)J.qlen, tlen = (int32_t)J.tlen;
    const int32_t K = (int32_t)J.k, dw = (K + 7) / 8;
    const uint32_t *zj = z + J.z_off;
    uint32_t *o = ops + J.ops_off;
    uint32_t cnt = 0, last = 0;  // `last` is the open run, not yet stored
    int32_t i = tlen - 1;
    int32_t c = min(tlen - 1 + w + 1, qlen) - 1;
    uint32_t which = 0;
    auto push = [&](uint32_t op, uint32_t len) {
        if (cnt > 0 && (last & 0xfu) == op) {
            last += len << 4;
        } else {
            if (cnt > 0) o[cnt - 1] = last;
            last = (len << 4) | op;
            ++cnt;
        }
    };
    while (i >= 0 && c >= 0) {
        const int32_t beg = max(i - w, 0);
        const int32_t off = c - beg;
        if (off < 0 || off >= 16 * K) {  // off the stored band: only without a path inside it
            cnt = kTracebackOffBand;
            break;
        }
        const int32_t g = off / K, kk = off - g * K;
        const int32_t m = kk >> 3, nib = min(8, K - 8 * m) - 1 - (kk & 7);
        const uint32_t code = (zj[code_index(i, g, m, dw)] >> (4 * nib)) & 15u;
        const uint32_t src = (code & 2u) ? 2u : (code & 1u);  // 0 M, 1 E, 2 F
        which = which == 0 ? src : (which == 1 ? (code >> 2) & 1u : ((code >> 3) & 1u) * 2u);
        if (which == 0) {
            push(0, 1);
            --i;
            --c;
        } else if (which == 1) {
            push(2, 1);
            --i;
        } else {
            push(1, 1);
            --c;
        }
    }
    if (cnt == kTracebackOffBand) {
        n_ops[k] = cnt;
        return;
    }
    if (i >= 0) push(2, (uint32_t)(i + 1));
    if (c >= 0) push(1, (uint32_t)(c + 1));
    if (cnt > 0) o[cnt - 1] = last;
    n_ops[k] = cnt;
}

// ---- long queries: one job per thread -------------------------------------------------------
// Queries beyond the group kernel's LDS columns (more than 512 bases, or a band wider than its
// widest K) run the same recurrence one job per thread, with H and E over absolute columns and one
// code byte per band cell in the job's region of z, and then walk their own codes back, leaving
// the runs back to front at ops + ops_off as glb_traceback_kernel does.  A code byte holds in bits
// 0-1 how H was reached (0 M, 1 E, 2 F; ties M over E over F), in bit 2 that the deletion
// continues and in bit 5 that the insertion continues, so that (code >> 2*state) & 3 is the walk's
// next state from state 0 (H), 1 (in a deletion) or 2 (in an insertion), sw.rs:650-688.
// Region of a job in z: (qlen + 2) {H, E} pairs, then tlen rows of min(qlen, 2w + 1) code bytes.
__global__ __launch_bounds__(64) void glb_plain_kernel(const uint8_t *__restrict__ arena,
                                                       const uint8_t *__restrict__ ref,
                                                       const GlobalJob *__restrict__ jobs,
                                                       const uint32_t *__restrict__ order, uint32_t count,
                                                       GlobalScoring sc, uint32_t *__restrict__ z,
                                                       uint32_t *__restrict__ ops, uint32_t *__restrict__ n_ops,
                                                       int32_t *__restrict__ score) {
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= count) return;
    const uint32_t idx = order[slot];
    const GlobalJob J = jobs[idx];
    const uint8_t *q = arena + J.q_off;
    const TargetView t = target_view(arena, ref, J.t_off, J.tlen);
    const int32_t qlen = (int32_t)J.qlen, tlen = (int32_t)J.tlen, w = max(J.w, 0);
    const int32_t ncol = min(qlen, 2 * w + 1);
    const int32_t oe_del = sc.o_del + sc.e_del, oe_ins = sc.o_ins + sc.e_ins;
    int2 *eh = (int2 *)(z + J.z_off);
    uint8_t *codes = (uint8_t *)(eh + qlen + 2);
    // row -1 (sw.rs:614-621): a leading insertion within the band, nothing open
    for (int32_t j = 0; j <= qlen; ++j)
        eh[j] = make_int2(j == 0 ? 0 : (j <= w ? -(sc.o_ins + sc.e_ins * j) : kMinusInf), kMinusInf);
    for (int32_t i = 0; i < tlen; ++i) {
        const int32_t beg = max(i - w, 0), end = min(i + w + 1, qlen);
        const int8_t *row = sc.mat + 5 * (int32_t)t[i];
        uint8_t *zr = codes + (size_t)i * (size_t)ncol;  // cell (i, j) at zr[j - beg]
        int32_t h_left = beg == 0 ? -(sc.o_del + sc.e_del * (i + 1)) : kMinusInf;  // H(i, j-1)
        int32_t f = kMinusInf;
        for (int32_t j = beg; j < end; ++j) {
            const int2 x = eh[j];  // H(i-1, j-1) and E(i, j)
            const int32_t m = x.x + row[q[j]];
            uint32_t code = m < x.y ? 1u : 0u;
            int32_t h = max(m, x.y);
            if (h < f) {
                code = 2u;
                h = f;
            }
            const int32_t e_ext = x.y - sc.e_del, e_open = m - oe_del;
            const int32_t f_ext = f - sc.e_ins, f_open = m - oe_ins;
            if (e_ext > e_open) code |= 4u;
            if (f_ext > f_open) code |= 32u;
            eh[j] = make_int2(h_left, max(e_ext, e_open));  // H(i, j-1): the next row's diagonal
            f = max(f_ext, f_open);
            h_left = h;
            zr[j - beg] = (uint8_t)code;
        }
        eh[end] = make_int2(h_left, kMinusInf);
    }
    score[idx] = eh[qlen].x;
    // the walk back from the last cell of the band (sw.rs:650-688), runs merged as push_cigar does
    uint32_t *o = ops + J.ops_off;
    uint32_t cnt = 0, last = 0;  // `last` is the open run, not yet stored
    auto push = [&](uint32_t op, uint32_t len) {
        if (cnt > 0 && (last & 0xfu) == op) {
            last += len << 4;
        } else {
            if (cnt > 0) o[cnt - 1] = last;
            last = (len << 4) | op;
            ++cnt;
        }
    };
    int32_t i = tlen - 1, c = min(tlen - 1 + w + 1, qlen) - 1;
    uint32_t state = 0;
    while (i >= 0 && c >= 0) {
        const int32_t off = c - max(i - w, 0);
        if (off < 0 || off >= ncol) {  // off the band: only without a path inside it
            n_ops[idx] = kTracebackOffBand;
            return;
        }
        const uint32_t code = codes[(size_t)i * (size_t)ncol + (size_t)off];
        state = (code >> (2 * state)) & 3u;
        if (state == 0) {
            push(0, 1);
            --i;
            --c;
        } else if (state == 1) {
            push(2, 1);
            --i;
        } else {
            push(1, 1);
            --c;
        }
    }
    if (i >= 0) push(2, (uint32_t)(i + 1));
    if (c >= 0) push(1, (uint32_t)(c + 1));
    if (cnt > 0) o[cnt - 1] = last;
    n_ops[idx] = cnt;
}

__global__ __launch_bounds__(256) void glb_cigar_gather_kernel(const GlobalJob *__restrict__ jobs, uint32_t n,
                                                               const uint32_t *__restrict__ ops,
                                                               const uint32_t *__restrict__ n_ops,
                                                               const uint64_t *__restrict__ dst_off,
                                                               uint32_t *__restrict__ dst) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t *src = ops + jobs[k].ops_off;
    const uint32_t m = n_ops[k];
    uint32_t *d = dst + dst_off[k];
    for (uint32_t r = 0; r < m; ++r) d[r] = src[m - 1 - r];
}

template <int K, bool PACKED>
int glb_occ(int qcap) {
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, glb_group_kernel<K, PACKED>, 256,
                                                     glb_block_lds_bytes(qcap, K, PACKED)) != hipSuccess)
        blocks = 1;
    return blocks > 0 ? blocks : 1;
}

template <int K, bool PACKED>
hipError_t launch_glb_group_k(hipStream_t st, const uint8_t *arena, const uint8_t *ref, const GlobalJob *jobs,
                              const uint32_t *order, uint32_t count, uint32_t *head, const GlobalScoring &sc,
                              int qcap, uint32_t *z, int32_t *score, int grid) {
    hipLaunchKernelGGL((glb_group_kernel<K, PACKED>), dim3(grid), dim3(256), glb_block_lds_bytes(qcap, K, PACKED), st,
                       arena, ref, jobs, order, count, head, sc, qcap, z, score);
    return hipGetLastError();
}

}  // namespace

bool glb_scores_packable(const GlobalScoring &sc) {
    for (int k = 0; k < 25; ++k)
        if (sc.mat[k] < -32 || sc.mat[k] > 31) return false;
    return true;
}

size_t glb_group_lds_bytes(int K, int qcap, bool packed) { return glb_block_lds_bytes(qcap, K, packed); }

int glb_group_blocks_per_cu(int K, int qcap, bool packed) {
    switch (K) {
#define BWG_K(k) case k: return packed ? glb_occ<k, true>(qcap) : glb_occ<k, false>(qcap);
        BWG_EXT_GROUP_KS
#undef BWG_K
    default: return 1;
    }
}

hipError_t launch_glb_group(hipStream_t st, int K, bool packed, const uint8_t *arena, const uint8_t *ref,
                            const GlobalJob *jobs, const uint32_t *order, uint32_t count, uint32_t *head,
                            const GlobalScoring &sc, int qcap, uint32_t *z, int32_t *score, int grid) {
    if (count == 0) return hipSuccess;
    switch (K) {
#define BWG_K(k)                                                                                                    \
    case k:                                                                                                         \
        return packed ? launch_glb_group_k<k, true>(st, arena, ref, jobs, order, count, head, sc, qcap, z, score, grid) \
                      : launch_glb_group_k<k, false>(st, arena, ref, jobs, order, count, head, sc, qcap, z, score, grid);
        BWG_EXT_GROUP_KS
#undef BWG_K
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_glb_traceback(hipStream_t st, const GlobalJob *jobs, uint32_t n, const uint32_t *z, uint32_t *ops,
                                uint32_t *n_ops) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(glb_traceback_kernel, dim3((n + 255) / 256), dim3(256), 0, st, jobs, n, z, ops, n_ops);
    return hipGetLastError();
}

hipError_t launch_glb_plain(hipStream_t st, const uint8_t *arena, const uint8_t *ref, const GlobalJob *jobs,
                            const uint32_t *order, uint32_t count, const GlobalScoring &sc, uint32_t *z,
                            uint32_t *ops, uint32_t *n_ops, int32_t *score) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(glb_plain_kernel, dim3((count + 63) / 64), dim3(64), 0, st, arena, ref, jobs, order, count, sc,
                       z, ops, n_ops, score);
    return hipGetLastError();
}

hipError_t launch_glb_cigar_gather(hipStream_t st, const GlobalJob *jobs, uint32_t n, const uint32_t *ops,
                                   const uint32_t *n_ops, const uint64_t *dst_off, uint32_t *dst) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(glb_cigar_gather_kernel, dim3((n + 255) / 256), dim3(256), 0, st, jobs, n, ops, n_ops,
                       dst_off, dst);
    return hipGetLastError();
}

}  // namespace bwg
