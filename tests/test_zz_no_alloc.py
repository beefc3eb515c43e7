"""GPU test of the allocation-free extension entry point, bwg_extend_batch_noalloc.

It is called through the C ABI and compared field by field with bwg_extend_batch on the same jobs,
on two calls in a row, the second with fewer jobs (the device scratch keeps the first call's
contents, so stale queue heads, counters or results would show).  The calls must not change the
context's device pool, and a batch beyond the scratch is rejected with BWG_E2BIG.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import bwg

pytestmark = pytest.mark.gpu

MAT = bwg.default_mat()


def _noalloc(ctx, batch, params):
    L = bwg.library()
    out = np.zeros(len(batch), dtype=bwg.EXT_DTYPE)
    n, ap, al, qo, ql, to, tl = batch._args()
    rc = L.bwg_extend_batch_noalloc(ctx.h, C.byref(params), n, ap, al, qo, ql, to, tl, batch.h0.ctypes.data,
                                    out.ctypes.data)
    return rc, out


def test_noalloc_matches_extend_batch_twice(gpu_ctx):
    print("noalloc: two calls, second smaller", file=sys.stderr, flush=True)
    gpu_ctx.set_form(bwg.FORM_BATCHED)
    p = bwg.ext_params(5, MAT, 6, 1, 6, 1, 100, 5, 100)
    big = bwg.generate(bwg.WORKLOAD_EXT_TAIL, 60_000, seed=31)
    small = bwg.generate(bwg.WORKLOAD_EXT_SMALL, 7_000, seed=32)
    # a batch with long-query jobs that the plain kernel takes (its state in the scratch too)
    rng = np.random.default_rng(3)
    qs = [rng.integers(0, 4, ql).astype(np.uint8) for ql in (9000, 600, 30, 140)]
    ts = [np.concatenate([q[:min(len(q), 500)], rng.integers(0, 4, 100).astype(np.uint8)]) for q in qs]
    odd = bwg.JobBatch.from_pairs(qs, ts, [40, 50, 60, 70])
    want_big = gpu_ctx.extend_arrays(big, p)
    want_small = gpu_ctx.extend_arrays(small, p)
    want_odd = gpu_ctx.extend_arrays(odd, p)
    # the no-alloc calls below must leave the context's device pool untouched
    in_use, cached = gpu_ctx.stat(bwg.STAT_POOL_IN_USE), gpu_ctx.stat(bwg.STAT_POOL_CACHED)
    launches = gpu_ctx.stat(bwg.STAT_LAUNCHES)
    rc, got = _noalloc(gpu_ctx, big, p)
    assert rc == 0, bwg.library().bwg_strerror(rc)
    assert np.array_equal(got, want_big)
    rc, got = _noalloc(gpu_ctx, small, p)
    assert rc == 0, bwg.library().bwg_strerror(rc)
    assert np.array_equal(got, want_small)
    rc, got = _noalloc(gpu_ctx, odd, p)
    assert rc == 0 and np.array_equal(got, want_odd)
    assert gpu_ctx.stat(bwg.STAT_LAUNCHES) > launches
    assert gpu_ctx.stat(bwg.STAT_POOL_IN_USE) == in_use
    assert gpu_ctx.stat(bwg.STAT_POOL_CACHED) == cached


def test_noalloc_rejects_batches_beyond_the_scratch(gpu_ctx):
    p = bwg.ext_params(5, MAT, 6, 1, 6, 1, 100, 5, 100)
    # 70 MB of arena: more than the 64 MiB scratch holds
    n = 1000
    q = np.zeros(35_000, dtype=np.uint8)
    batch = bwg.JobBatch.from_pairs([q] * n, [q[:35_000]] * n, np.full(n, 30))
    rc, _ = _noalloc(gpu_ctx, batch, p)
    assert rc == bwg.BWG_E2BIG
    rc, _ = _noalloc(gpu_ctx, bwg.generate(bwg.WORKLOAD_EXT_SMALL, 10, seed=1), p)
    assert rc == 0
