"""GPU tests of the host-side split of a large synchronous call.

A call of many jobs validates its inputs, builds its descriptors, classifies and bins its jobs over
several host threads, and uploads a large arena from a helper thread while it plans.  None of that
may change a result, the binning, or which error a bad call returns: the error is the one the
first failing job (in index order) gives, after the arena-wide checks, as with one thread.
"""
import sys

import numpy as np
import pytest

import bwg
import oracle_binding as ob

pytestmark = pytest.mark.gpu

MAT = bwg.default_mat()
EXT = (6, 1, 6, 1, 100, 5, 100)
KSW = (6, 1, 6, 1, 20, 1)


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


@pytest.fixture
def threads_option(gpu_ctx):
    yield lambda v: gpu_ctx.set_option(bwg.OPT_HOST_THREADS, v)
    gpu_ctx.set_option(bwg.OPT_HOST_THREADS, 0)


def test_extension_results_and_bins_do_not_depend_on_host_threads(gpu_ctx, threads_option):
    _say("host threads: 200k extension jobs with 1, 16 and automatic host threads")
    gpu_ctx.set_form(bwg.FORM_BATCHED)
    p = bwg.ext_params(5, MAT, *EXT)
    batch = bwg.generate(bwg.WORKLOAD_EXT_TAIL, 200_000, seed=91)
    want = ob.extend_batch(batch, MAT, *EXT, nthreads=8)
    outs, bins = [], []
    for v in (1, 16, 0):
        threads_option(v)
        outs.append(gpu_ctx.extend_arrays(batch, p))
        assert gpu_ctx.phases()["total"] > 0
        st = gpu_ctx.stage(batch, p)
        st.run()
        bins.append(st.job_bins())
        assert np.array_equal(st.results(), want)
        st.close()
    for o in outs:
        assert np.array_equal(o, want)
    assert np.array_equal(bins[0], bins[1]) and np.array_equal(bins[0], bins[2])


def test_rescue_results_do_not_depend_on_host_threads(gpu_ctx, threads_option):
    _say("host threads: 20k rescue jobs with 1 and 16 host threads")
    gpu_ctx.set_form(bwg.FORM_BATCHED)
    kp = bwg.ksw_params(5, MAT, *KSW)
    batch = bwg.generate(bwg.WORKLOAD_RESCUE, 20_000, seed=92)
    want = ob.align2_batch(batch, MAT, *KSW)
    for v in (1, 16):
        threads_option(v)
        assert np.array_equal(gpu_ctx.align2_arrays(batch, kp), want)


def _code(gpu_ctx, batch, p):
    try:
        gpu_ctx.extend_arrays(batch, p)
    except bwg.BwgError as e:
        return e.code
    return 0


def test_first_failing_job_decides_the_error(gpu_ctx, threads_option):
    """Rejected before anything is staged: no kernel runs for these calls."""
    _say("host threads: error precedence over 50k jobs")
    p = bwg.ext_params(5, MAT, *EXT)
    base = bwg.generate(bwg.WORKLOAD_EXT_TAIL, 50_000, seed=93)

    def variant(einval_at=None, e2big_at=None, bad_bounds_at=None, bad_code_at=None):
        h0 = base.h0.copy()
        q_off = base.q_off.copy()
        arena = base.arena
        if einval_at is not None:
            h0[einval_at] = 0
        if e2big_at is not None:
            h0[e2big_at] = (1 << 30) - 5
        if bad_bounds_at is not None:
            q_off[bad_bounds_at] = len(arena)
        if bad_code_at is not None:
            arena = arena.copy()
            arena[int(base.q_off[bad_code_at])] = 7
        return bwg.JobBatch(arena, q_off, base.q_len, base.t_off, base.t_len, h0)

    for v in (1, 16, 0):
        threads_option(v)
        assert _code(gpu_ctx, variant(einval_at=40_000, e2big_at=45_000), p) == bwg.BWG_EINVAL
        assert _code(gpu_ctx, variant(einval_at=45_000, e2big_at=40_000), p) == bwg.BWG_E2BIG
        assert _code(gpu_ctx, variant(einval_at=3, e2big_at=2), p) == bwg.BWG_E2BIG
        # the arena-wide checks come first: bounds, then base codes, then the per-job tests
        assert _code(gpu_ctx, variant(e2big_at=100, bad_bounds_at=49_999), p) == bwg.BWG_EINVAL
        assert _code(gpu_ctx, variant(e2big_at=100, bad_code_at=49_000), p) == bwg.BWG_EINVAL
    # and the valid batch still runs
    threads_option(16)
    assert np.array_equal(gpu_ctx.extend_arrays(base, p), ob.extend_batch(base, MAT, *EXT, nthreads=8))
