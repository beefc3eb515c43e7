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
//    second pass.  Every evaluated cell leaves a 4-bit traceback code in HBM: bits 0-1 how H was
//    reached (0 M, 1 E, 2 F; ties M over E over F), bit 2 "the deletion continues", bit 3 "the
//    insertion continues", the same information as the reference's byte f<<4 | e<<2 | h.  The
//    codes of a row are stored lane-major: lane g's K codes in (K+7)/8 dwords at
//    (i*16 + g) * dwords, so each lane writes whole dwords.
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
__host__ __device__ __forceinline__ int glb_plane_bytes(int qcap) { return 5 * qcap; }
__host__ __device__ __forceinline__ size_t glb_block_lds_bytes(int qcap, int K) {
    return (size_t)kGlbGroupsPerBlock * (glb_eh_bytes(qcap) + glb_plane_bytes(qcap)) + 16 * K * 8 + 64;
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

template <int K>
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
    int32_t *eh = (int32_t *)((char *)lds + gid * glb_eh_bytes(qcap));
    int8_t *planes = (int8_t *)lds + kGlbGroupsPerBlock * glb_eh_bytes(qcap) + gid * glb_plane_bytes(qcap);
    const int32_t oe_del = sc.o_del + sc.e_del, oe_ins = sc.o_ins + sc.e_ins;
    const int32_t e_del = sc.e_del, e_ins = sc.e_ins;
    const int32_t Ke = K * e_ins;

    bool active = false, exhausted = false;
    TargetView t = {nullptr, 1};
    int32_t qlen = 0, tlen = 0, w = 0, i = 0, tcur = 0;
    uint32_t idx = 0;
    uint32_t *zj = nullptr;
    for (;;) {
        if (!active && !exhausted) {
            uint32_t slot = 0;
            if (g == 0) slot = atomicAdd(head, 1u);
            slot = __shfl(slot, lane & ~15, 64);
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
#pragma unroll
                    for (int c = 0; c < 5; ++c) planes[c * qcap + j] = sc.mat[c * 5 + qc];
                }
                i = 0;
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
        const int32_t h1_init = beg == 0 ? -(sc.o_del + sc.e_del * (i + 1)) : kMinusInf;
        const int32_t tnext = i + 1 < tlen ? t[i + 1] : 0;
        const int32_t base = beg + g * K;
        const int32_t nval = min(max(end - base, 0), K);
        const int2 *ehb = (const int2 *)eh + base;
        const int8_t *sb = planes + tcur * qcap + base;
        int2 v[K];
        int32_t s[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            v[k] = ehb[k];
            s[k] = sb[k];
        }
        // pass 1: M and the lane's own F chain from the identity
        int32_t mm[K], gg[K];
        int32_t grun = kNeg;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            mm[k] = v[k].x + s[k];
            gg[k] = grun;
            grun = max(grun - e_ins, mm[k] - oe_ins);
        }
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
            const int32_t f = max(f_in - k * e_ins, gg[k]);
            const int32_t e = v[k].y;
            uint32_t code = mm[k] < e ? 1u : 0u;
            int32_t h = max(mm[k], e);
            if (h < f) {
                code = 2u;
                h = f;
            }
            const int32_t td = mm[k] - oe_del, en = e - e_del;
            code |= en > td ? 4u : 0u;
            code |= f - e_ins > mm[k] - oe_ins ? 8u : 0u;
            zw[k / 8] |= code << (4 * (k % 8));
            if (k < nval) {
                ehw[k].y = max(en, td);  // E(i+1, j)
                ehw[k + 1].x = h;        // H(i, j), the diagonal of column j+1 next row
            }
        }
        if (g == 0) {
            eh[2 * beg] = h1_init;        // H(i, beg-1)
            eh[2 * end + 1] = kMinusInf;  // E at the band's end
        }
        if (nval > 0) {
            uint32_t *zr = zj + (size_t)(i * 16 + g) * DW;
#pragma unroll
            for (int m = 0; m < DW; ++m)
                if (8 * m < nval) zr[m] = zw[m];
        }
        wave_lds_sync();
        ++i;
        tcur = tnext;
        if (i >= tlen) {
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
        const uint32_t code = (zj[(size_t)(i * 16 + g) * dw + kk / 8] >> (4 * (kk % 8))) & 15u;
        which = which == 0 ? (code & 3u) : (which == 1 ? (code >> 2) & 1u : ((code >> 3) & 1u) * 2u);
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

template <int K>
int glb_occ(int qcap) {
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, glb_group_kernel<K>, 256,
                                                     glb_block_lds_bytes(qcap, K)) != hipSuccess)
        blocks = 1;
    return blocks > 0 ? blocks : 1;
}

}  // namespace

size_t glb_group_lds_bytes(int K, int qcap) { return glb_block_lds_bytes(qcap, K); }

int glb_group_blocks_per_cu(int K, int qcap) {
    switch (K) {
#define BWG_K(k) case k: return glb_occ<k>(qcap);
        BWG_EXT_GROUP_KS
#undef BWG_K
    default: return 1;
    }
}

hipError_t launch_glb_group(hipStream_t st, int K, const uint8_t *arena, const uint8_t *ref, const GlobalJob *jobs,
                            const uint32_t *order, uint32_t count, uint32_t *head, const GlobalScoring &sc,
                            int qcap, uint32_t *z, int32_t *score, int grid) {
    if (count == 0) return hipSuccess;
    switch (K) {
#define BWG_K(k)                                                                                        \
    case k:                                                                                             \
        hipLaunchKernelGGL(glb_group_kernel<k>, dim3(grid), dim3(256), glb_block_lds_bytes(qcap, k), st, \
                           arena, ref, jobs, order, count, head, sc, qcap, z, score);                   \
        return hipGetLastError();
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

hipError_t launch_glb_cigar_gather(hipStream_t st, const GlobalJob *jobs, uint32_t n, const uint32_t *ops,
                                   const uint32_t *n_ops, const uint64_t *dst_off, uint32_t *dst) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(glb_cigar_gather_kernel, dim3((n + 255) / 256), dim3(256), 0, st, jobs, n, ops, n_ops,
                       dst_off, dst);
    return hipGetLastError();
}

}  // namespace bwg
