"""GPU tests of the synchronous entry points at the reference's call shapes.

The reference calls SwBackend::extend_batch from rayon workers, one call per chunk
(crates/bwa-cli/src/cmd_mem.rs:2205-2216), and batched_ksw_align2 with about a hundred jobs per
call from par_chunks_mut (crates/bwa-mem/src/pe.rs:1090-1100).  Here 16 host threads call
bwg_extend_batch and bwg_ksw_align2_batch concurrently on one context (ctypes releases the GIL for
the duration of each call), while another thread changes options; every result must equal the
oracle's.  Also: jobs beyond the binned kernels' shapes (long queries and targets) routed to the
plain kernels, and the server's ticket release when poll and wait meet on one ticket.
"""
import sys
import threading

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


@pytest.fixture(scope="module")
def call_sets():
    """32 extension calls of 2000 long-tail jobs and 32 rescue calls of 100 jobs, each with its
    oracle results."""
    ext = bwg.generate(bwg.WORKLOAD_EXT_TAIL, 64_000, seed=71)
    resc = bwg.generate(bwg.WORKLOAD_RESCUE, 3_200, seed=72)
    ext_want = ob.extend_batch(ext, MAT, *EXT)
    resc_want = ob.align2_batch(resc, MAT, *KSW)
    ext_calls = [(ext.subset(np.arange(k * 2000, (k + 1) * 2000)), ext_want[k * 2000:(k + 1) * 2000])
                 for k in range(32)]
    resc_calls = [(resc.subset(np.arange(k * 100, (k + 1) * 100)), resc_want[k * 100:(k + 1) * 100])
                  for k in range(32)]
    return ext_calls, resc_calls


def test_sixteen_threads_one_context(gpu_ctx, call_sets):
    _say("sixteen_threads_one_context")
    ext_calls, resc_calls = call_sets
    gpu_ctx.set_form(bwg.FORM_BATCHED)
    ep = bwg.ext_params(5, MAT, *EXT)
    kp = bwg.ksw_params(5, MAT, *KSW)
    bad, errors = [], []
    stop = threading.Event()

    def worker(t):
        try:
            for r in range(4):
                k = (t * 4 + r) % 32
                if (t + r) % 2 == 0:
                    batch, want = ext_calls[k]
                    got = gpu_ctx.extend_arrays(batch, ep)
                else:
                    batch, want = resc_calls[k]
                    got = gpu_ctx.align2_arrays(batch, kp)
                if not np.array_equal(got, want):
                    bad.append((t, r, k))
        except Exception as e:  # surfaced below
            errors.append(repr(e))

    def toggler():
        # every mapping these options select is exact, so results must not move while they change
        v = 0
        while not stop.is_set():
            v ^= 1
            gpu_ctx.set_option(bwg.OPT_EXT_K_STEP, v)
            gpu_ctx.set_option(bwg.OPT_CONCURRENT_BINS, 1 - v)

    launches0 = gpu_ctx.stat(bwg.STAT_LAUNCHES)
    tog = threading.Thread(target=toggler)
    tog.start()
    th = [threading.Thread(target=worker, args=(t,)) for t in range(16)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    stop.set()
    tog.join()
    gpu_ctx.set_option(bwg.OPT_EXT_K_STEP, 0)
    gpu_ctx.set_option(bwg.OPT_CONCURRENT_BINS, 1)
    assert not errors, errors[:3]
    assert not bad, bad[:5]
    assert gpu_ctx.stat(bwg.STAT_LAUNCHES) > launches0


def test_lanes_option_and_single_lane(gpu_ctx, call_sets):
    """With one lane every call still completes (the others wait for it)."""
    _say("lanes_option")
    ext_calls, _ = call_sets
    ep = bwg.ext_params(5, MAT, *EXT)
    gpu_ctx.set_option(bwg.OPT_LANES, 1)
    try:
        out = [None] * 6

        def w(i):
            out[i] = gpu_ctx.extend_arrays(ext_calls[i][0], ep)

        th = [threading.Thread(target=w, args=(i,)) for i in range(6)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        for i in range(6):
            assert np.array_equal(out[i], ext_calls[i][1])
    finally:
        gpu_ctx.set_option(bwg.OPT_LANES, 4)
    with pytest.raises(bwg.BwgError):
        gpu_ctx.set_option(bwg.OPT_LANES, 0)


def _random_seq(rng, n):
    return rng.integers(0, 4, n).astype(np.uint8)


def _mutate(rng, s, sub=0.02):
    s = s.copy()
    m = rng.random(len(s)) < sub
    s[m] = (s[m] + rng.integers(1, 4, m.sum())) % 4
    return s


def test_long_jobs_run_in_plain_kernels(gpu_ctx):
    """Queries and targets beyond the binned kernels' shapes (qlen > 8191 columns, qlen > 512 with
    wide bands, rescue targets > 65535 rows, rescue queries padded beyond 256 columns) run on the
    device and equal the oracle, in a batch mixed with ordinary jobs."""
    _say("long_jobs_plain")
    rng = np.random.default_rng(5)
    qs, ts, h0 = [], [], []
    for ql, tl in ((9000, 9100), (5000, 5200), (700, 900), (40, 140), (20000, 300)):
        q = _random_seq(rng, ql)
        t = np.concatenate([_mutate(rng, q[:min(ql, tl)]), _random_seq(rng, max(0, tl - ql))])[:tl]
        qs.append(q)
        ts.append(t)
        h0.append(int(rng.integers(19, 200)))
    batch = bwg.JobBatch.from_pairs(qs, ts, h0)
    for form in (bwg.FORM_BATCHED, bwg.FORM_PLAIN):
        gpu_ctx.set_form(form)
        for w in (100, 5000):
            ep = bwg.ext_params(5, MAT, 6, 1, 6, 1, w, 5, 100)
            got = gpu_ctx.extend_arrays(batch, ep)
            want = ob.extend_batch(batch, MAT, 6, 1, 6, 1, w, 5, 100)
            assert np.array_equal(got, want), (form, w)
    gpu_ctx.set_form(bwg.FORM_BATCHED)
    # rescue: a 150 bp read in a 70k-row window (planted), a 5000 bp query, a 300 bp query
    qs, ts = [], []
    read = _random_seq(rng, 150)
    win = _random_seq(rng, 70_000)
    win[41_000:41_150] = _mutate(rng, read)
    qs.append(read)
    ts.append(win)
    q = _random_seq(rng, 5000)
    qs.append(q)
    ts.append(np.concatenate([_random_seq(rng, 300), _mutate(rng, q[:1500]), _random_seq(rng, 200)]))
    q = _random_seq(rng, 300)
    qs.append(q)
    ts.append(np.concatenate([_random_seq(rng, 500), _mutate(rng, q), _random_seq(rng, 400)]))
    qs.append(_random_seq(rng, 150))
    ts.append(_random_seq(rng, 900))
    batch = bwg.JobBatch.from_pairs(qs, ts)
    kp = bwg.ksw_params(5, MAT, *KSW)
    got = gpu_ctx.align2_arrays(batch, kp)
    want = ob.align2_batch(batch, MAT, *KSW)
    assert np.array_equal(got, want)
    assert got[0]["te"] >= 41_100


def test_size_limits_still_reject_overflowing_jobs(gpu_ctx):
    """Beyond 2^24 bases, or where an int32 score could overflow, the call fails loudly."""
    ep = bwg.ext_params(5, MAT, 6, 1, 6, 1, 100, 5, 100)
    batch = bwg.JobBatch.from_pairs([np.zeros(10, np.uint8)], [np.zeros(10, np.uint8)], [(1 << 30) - 5])
    with pytest.raises(bwg.BwgError) as e:
        gpu_ctx.extend_arrays(batch, ep)
    assert e.value.code == bwg.BWG_E2BIG


def test_server_poll_while_waiting(gpu_ctx):
    """A thread polls a ticket another thread waits on: the waiter gets the status and releases the
    ticket; polling never frees it under the waiter."""
    _say("server_poll_while_waiting")
    ep = bwg.ext_params(5, MAT, *EXT)
    batch = bwg.generate(bwg.WORKLOAD_EXT_TAIL, 20_000, seed=9)
    want = ob.extend_batch(batch, MAT, *EXT)
    srv = bwg.Server(gpu_ctx, hbm_budget_bytes=4 << 30, slots=2)
    try:
        for _ in range(5):
            out = np.zeros(len(batch), dtype=bwg.EXT_DTYPE)
            tk = srv.submit_into(batch, ep, out)
            res = {}

            def waiter():
                res["rc"] = srv.wait_rc(tk)

            th = threading.Thread(target=waiter)
            th.start()
            polled = []
            while th.is_alive():
                polled.append(srv.poll_rc(tk))
            th.join()
            assert res["rc"] == bwg.BWG_OK
            assert all(p in (bwg.BWG_EAGAIN, bwg.BWG_OK, bwg.BWG_EINVAL) for p in polled)
            assert srv.poll_rc(tk) == bwg.BWG_EINVAL  # released by the waiter
            assert np.array_equal(out, want)
    finally:
        srv.close()


def test_server_admission_counts_plain_rescue_state(gpu_ctx):
    """Chunks of long-query rescue jobs (plain bin: per-job {H,E} columns and row maxima in device
    memory) under a budget that admits about one chunk: the device bytes the context's batches
    actually held stay within the budget, and every chunk is exact."""
    _say("server_admission_plain_rescue")
    rng = np.random.default_rng(11)
    qs, ts = [], []
    for _ in range(60):
        q = _random_seq(rng, 600)
        qs.append(q)
        ts.append(np.concatenate([_random_seq(rng, 700), _mutate(rng, q), _random_seq(rng, 800)]))
    batch = bwg.JobBatch.from_pairs(qs, ts)
    kp = bwg.ksw_params(5, MAT, *KSW)
    want = ob.align2_batch(batch, MAT, *KSW)
    budget = 24 << 20
    srv = bwg.Server(gpu_ctx, hbm_budget_bytes=budget, slots=3)
    try:
        base = gpu_ctx.stat(bwg.STAT_POOL_IN_USE)
        gpu_ctx.set_option(bwg.OPT_POOL_PEAK_RESET, 1)
        tks = [srv.submit(batch, kp) for _ in range(6)]
        for tk in tks:
            assert np.array_equal(srv.wait(tk), want)
        cnt = srv.counters()
        assert cnt["peak_device_bytes"] <= budget
        assert cnt["peak_pool_bytes"] - base <= budget, (cnt["peak_pool_bytes"], base)
    finally:
        srv.close()
