// Mate-rescue windows on the device: the reference's matesw_collect (crates/bwa-mem/src/pe.rs:596-679)
// with its fetch bns_fetch_seq (pe.rs:211-269), depos / pos2rid (crates/bwa-index/src/bntseq.rs:237-243,
// 337-367) and mem_infer_dir (pe.rs:192-209), turned into job descriptors that point into the
// device-resident two-strand reference, so no window sequence is ever copied.
//
// The reference is held as 2*l_pac bytes: the forward strand, then its reverse complement, so
// position p of the doubled coordinate space is byte p (fmindex.rs:281-310 computes the same bases
// from the packed forward strand).
#include <hip/hip_runtime.h>

#include "bwg_batched_kernels.h"

namespace bwg {
namespace {

__device__ __forceinline__ uint8_t complement(uint8_t c) { return c < 4 ? (uint8_t)(3 - c) : c; }

__global__ void reverse_half_kernel(uint8_t *ref2, int64_t l_pac) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < l_pac) ref2[l_pac + k] = complement(ref2[l_pac - 1 - k]);
}

// The mate's reverse complement (pe.rs:636-641), one block per anchor.
__global__ void mate_revcomp_kernel(uint8_t *arena, const MateAnchor *__restrict__ anchors, uint32_t n) {
    const uint32_t k = blockIdx.x;
    if (k >= n) return;
    const MateAnchor a = anchors[k];
    const uint8_t *src = arena + a.ms_off;
    uint8_t *dst = arena + a.rc_off;
    for (uint32_t j = threadIdx.x; j < a.l_ms; j += blockDim.x) dst[j] = complement(src[a.l_ms - 1 - j]);
}

// pos2rid: binary search over contig indices; the index of the last probe is the answer, as in the
// reference (a position before the first contig therefore reports contig 0).
__device__ int32_t pos2rid(int64_t pos_f, int64_t l_pac, const int64_t *__restrict__ off, int32_t n) {
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

// One thread per anchor: the skip test, then the four orientations' windows and job descriptors.
__global__ __launch_bounds__(256) void matesw_window_kernel(
    const MateAnchor *__restrict__ anchors, uint32_t n, const int64_t *__restrict__ ma_rb, MateRescueParams mp,
    const int64_t *__restrict__ contig_off, const int32_t *__restrict__ contig_len, KswJob *__restrict__ jobs,
    const int32_t *__restrict__ plain_slot, KswJob *__restrict__ plain_jobs, bwg_rescue_window *__restrict__ win) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const MateAnchor a = anchors[k];
    const int64_t l_pac = mp.l_pac, l2 = l_pac << 1;
    int skip[4];
    for (int r = 0; r < 4; ++r) skip[r] = mp.failed[r] ? 1 : 0;
    for (uint64_t e = a.ma_begin; e < a.ma_end; ++e) {  // mem_infer_dir(l_pac, a.rb, existing.rb)
        const int64_t b1 = a.rb, b2 = ma_rb[e];
        const int r1 = b1 >= l_pac, r2 = b2 >= l_pac;
        const int64_t p2 = r1 == r2 ? b2 : l2 - 1 - b2;
        const int64_t dist = p2 > b1 ? p2 - b1 : b1 - p2;
        const int dir = (r1 == r2 ? 0 : 1) ^ (p2 > b1 ? 0 : 3);
        if (dist >= mp.low[dir] && dist <= mp.high[dir]) skip[dir] = 1;
    }
    const bool none = skip[0] + skip[1] + skip[2] + skip[3] == 4;
    const int64_t l_ms = a.l_ms;
    for (int r = 0; r < 4; ++r) {
        const uint32_t j = 4 * k + (uint32_t)r;
        bwg_rescue_window w;
        w.rb = w.re = 0;
        w.rid = -1;
        w.flags = skip[r] ? BWG_WIN_SKIPPED : 0;
        const bool is_rev = (r >> 1) != (r & 1), is_larger = (r >> 1) == 0;
        if (is_rev) w.flags |= BWG_WIN_REV;
        bool valid = false;
        if (!skip[r] && !none) {
            const int64_t lo = mp.low[r], hi = mp.high[r];
            int64_t rb0, re0;
            if (!is_rev) {
                rb0 = is_larger ? a.rb + lo : a.rb - hi;
                re0 = (is_larger ? a.rb + hi : a.rb - lo) + l_ms;
            } else {
                rb0 = (is_larger ? a.rb + lo : a.rb - hi) - l_ms;
                re0 = is_larger ? a.rb + hi : a.rb - lo;
            }
            rb0 = rb0 > 0 ? rb0 : 0;
            re0 = re0 < l2 ? re0 : l2;
            if (rb0 < re0) {
                // bns_fetch_seq(rb0, (rb0 + re0) >> 1, re0)
                int64_t rb = rb0, re = re0;
                const int64_t mid = (rb0 + re0) >> 1;
                const bool strand_rev = mid >= l_pac;
                const int64_t fwd_mid = strand_rev ? l2 - 1 - mid : mid;
                const int32_t rid = pos2rid(fwd_mid, l_pac, contig_off, mp.n_contigs);
                int64_t far_beg = 0, far_end = l2;
                if (rid >= 0) {
                    far_beg = contig_off[rid];
                    far_end = far_beg + contig_len[rid];
                }
                if (strand_rev) {
                    const int64_t ob = far_beg;
                    far_beg = l2 - far_end;
                    far_end = l2 - ob;
                }
                rb = rb > far_beg ? rb : far_beg;
                re = re < far_end ? re : far_end;
                w.rb = rb;
                w.re = re;
                w.rid = rid;
                valid = a.rid == rid && re - rb >= mp.min_seed_len;
            }
        }
        if (valid) w.flags |= BWG_WIN_VALID;
        win[j] = w;
        // the descriptor: an invalid window becomes an empty target (default result)
        KswJob d = jobs[j];
        d.q_off = is_rev ? a.rc_off : a.ms_off;
        d.qlen = a.l_ms;
        d.t_off = kTFromRef | (uint64_t)(valid ? w.rb : 0);
        d.tlen = valid ? (uint32_t)(w.re - w.rb) : 0u;
        jobs[j] = d;
        if (plain_slot && plain_slot[j] >= 0) {
            KswJob pd = plain_jobs[plain_slot[j]];
            pd.q_off = d.q_off;
            pd.qlen = d.qlen;
            pd.t_off = d.t_off;
            pd.tlen = d.tlen;
            plain_jobs[plain_slot[j]] = pd;
        }
    }
}

}  // namespace

hipError_t launch_reverse_half(hipStream_t st, uint8_t *ref2, int64_t l_pac) {
    if (l_pac <= 0) return hipSuccess;
    hipLaunchKernelGGL(reverse_half_kernel, dim3((unsigned)((l_pac + 255) / 256)), dim3(256), 0, st, ref2, l_pac);
    return hipGetLastError();
}

hipError_t launch_mate_revcomp(hipStream_t st, uint8_t *arena, const MateAnchor *anchors, uint32_t n) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(mate_revcomp_kernel, dim3(n), dim3(128), 0, st, arena, anchors, n);
    return hipGetLastError();
}

hipError_t launch_matesw_windows(hipStream_t st, const MateAnchor *anchors, uint32_t n, const int64_t *ma_rb,
                                 const MateRescueParams &mp, const int64_t *contig_off, const int32_t *contig_len,
                                 KswJob *jobs, const int32_t *plain_slot, KswJob *plain_jobs,
                                 bwg_rescue_window *win) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(matesw_window_kernel, dim3((n + 255) / 256), dim3(256), 0, st, anchors, n, ma_rb, mp,
                       contig_off, contig_len, jobs, plain_slot, plain_jobs, win);
    return hipGetLastError();
}

}  // namespace bwg
