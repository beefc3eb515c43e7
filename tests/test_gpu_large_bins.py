"""Bins with more jobs than their persistent grid has 16-lane groups: the group kernels then deal
most job slots round-robin and take only the grid's last slots from the queue counter.  Every
output is checked against the oracle.  Small batches never reach that path (their grid is sized
to the bin), so these batches are built to fill one bin each."""
import sys

import numpy as np
import pytest

import bwg
import oracle_binding as ob

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


def _pairs(n, qlen, tlen, seed, n_frac=0.0):
    """n query/target pairs: the target starts with the query under 2% substitutions and runs on
    at random; n_frac of the queries get one N base."""
    rng = np.random.default_rng(seed)
    qs, ts = [], []
    for k in range(n):
        q = rng.integers(0, 4, qlen).astype(np.uint8)
        t = rng.integers(0, 4, tlen).astype(np.uint8)
        t[:qlen] = q
        m = rng.random(qlen) < 0.02
        t[:qlen][m] = (t[:qlen][m] + 1) % 4
        if rng.random() < n_frac:
            q[rng.integers(0, qlen)] = 4
        qs.append(q)
        ts.append(t)
    return qs, ts


def _first_bad(got, want):
    g, w = got.view(np.int32).reshape(len(got), -1), want.view(np.int32).reshape(len(want), -1)
    bad = np.nonzero((g != w).any(axis=1))[0]
    return None if bad.size == 0 else (int(bad[0]), g[bad[0]].tolist(), w[bad[0]].tolist())


def test_extension_one_full_bin(gpu_ctx):
    # 60k jobs of one window class (qlen 100, tlen 200 at w = 100), fast path off so that every
    # job reaches the group kernel
    _say("large bins: extension 60000 jobs, qlen 100")
    qs, ts = _pairs(60000, 100, 200, seed=91)
    rng = np.random.default_rng(92)
    batch = bwg.JobBatch.from_pairs(qs, ts, h0=rng.integers(19, 120, len(qs)).astype(np.int32))
    mat = bwg.default_mat()
    gpu_ctx.set_form(bwg.FORM_BATCHED)
    gpu_ctx.set_option(bwg.OPT_UNGAPPED, 0)
    try:
        got = gpu_ctx.extend_arrays(batch, bwg.ext_params(5, mat, 6, 1, 6, 1, 100, 5, 100))
    finally:
        gpu_ctx.set_option(bwg.OPT_UNGAPPED, 1)
    want = ob.extend_batch(batch, mat, 6, 1, 6, 1, 100, 5, 100, nthreads=16)
    assert _first_bad(got, want) is None, _first_bad(got, want)


def test_extension_short_queries_full_bin(gpu_ctx):
    # 600k 24-base queries (one job per lane): more jobs than the thread kernel's grid has lanes,
    # fast path off; the arena is built in one piece
    _say("large bins: extension 600000 jobs, qlen 24")
    n, L, T = 600000, 24, 124
    rng = np.random.default_rng(94)
    q = rng.integers(0, 4, (n, L)).astype(np.uint8)
    t = rng.integers(0, 4, (n, T)).astype(np.uint8)
    t[:, :L] = np.where(rng.random((n, L)) < 0.03, (q + 1) % 4, q)
    arena = np.concatenate([q, t], axis=1).reshape(-1)
    base = np.arange(n, dtype=np.uint64) * (L + T)
    batch = bwg.JobBatch(arena, base, np.full(n, L), base + L, np.full(n, T),
                         h0=rng.integers(19, 60, n).astype(np.int32))
    mat = bwg.default_mat()
    gpu_ctx.set_form(bwg.FORM_BATCHED)
    gpu_ctx.set_option(bwg.OPT_UNGAPPED, 0)
    try:
        got = gpu_ctx.extend_arrays(batch, bwg.ext_params(5, mat, 6, 1, 6, 1, 100, 5, 100))
    finally:
        gpu_ctx.set_option(bwg.OPT_UNGAPPED, 1)
    want = ob.extend_batch(batch, mat, 6, 1, 6, 1, 100, 5, 100, nthreads=16)
    assert _first_bad(got, want) is None, _first_bad(got, want)


def test_rescue_one_full_bin(gpu_ctx):
    # 40k 150-base mates against 180-base windows: one packed bin, whose reverse pass deals its
    # slots; 2% of the mates carry an N base
    _say("large bins: rescue 40000 jobs")
    qs, ts = _pairs(40000, 150, 180, seed=93, n_frac=0.02)
    batch = bwg.JobBatch.from_pairs(qs, ts)
    mat = bwg.default_mat()
    gpu_ctx.set_form(bwg.FORM_BATCHED)
    got = gpu_ctx.align2_arrays(batch, bwg.ksw_params(5, mat, 6, 1, 6, 1, 20, 1))
    want = ob.align2_batch(batch, mat, 6, 1, 6, 1, 20, 1, nthreads=16)
    assert _first_bad(got, want) is None, _first_bad(got, want)
    assert (got["qb"] >= 0).sum() > 30000  # most jobs ran the reverse pass


def test_global_one_full_bin(gpu_ctx):
    # 40k 100-base regions with the reference's band: one bin of the global group kernel
    _say("large bins: global 40000 regions")
    batch, w = bwg.generate_global_jobs(40000, seed=95, read_lens=(100,))
    mat = bwg.default_mat()
    gpu_ctx.set_form(bwg.FORM_BATCHED)
    got_s, got_n, got_c = gpu_ctx.global_arrays(batch, bwg.global_params(5, mat, 6, 1, 6, 1), w)
    want_s, want_n, want_c = ob.global_batch(batch, w, mat, 6, 1, 6, 1)
    assert np.array_equal(got_s, want_s)
    assert np.array_equal(got_n, want_n.astype(np.uint32))
    assert np.array_equal(got_c, want_c)


def test_short_query_bin_sized_from_another_batch_share():
    # After the ungapped pre-pass the one-job-per-lane bin's grid is sized from the share of its
    # jobs the pre-pass declined, learned on earlier runs (per context and mapping).  Batch A's
    # short queries are exact copies of their targets' starts: the walk accepts them all and the
    # context learns a share near 0.  Batch B's short queries sit against random targets, so the
    # walk declines nearly all of them; B's first run is sized from A's share, with a grid of a few
    # wavefronts for 100k declined jobs, and the persistent loop must take every other job from
    # the queue counter.  Both batches' runs (eager, then the replayed graph sized from B's own
    # share) equal the oracle.  Each batch also holds longer queries, so that its bins run
    # concurrently (the short bin's grid is resized only then).
    _say("short-query bin sized from another batch's declined share")
    rng = np.random.default_rng(95)
    mat = bwg.default_mat()
    params = bwg.ext_params(5, mat, 6, 1, 6, 1, 100, 5, 100)
    longq, longt = _pairs(4000, 120, 240, seed=96)

    def build(n_short, exact):
        qs, ts = list(longq), list(longt)
        for _ in range(n_short):
            q = rng.integers(0, 4, int(rng.integers(12, 33))).astype(np.uint8)
            t = rng.integers(0, 4, len(q) + int(rng.integers(0, 40))).astype(np.uint8)
            if exact:
                t[:len(q)] = q
            qs.append(q)
            ts.append(t)
        return bwg.JobBatch.from_pairs(qs, ts, h0=rng.integers(19, 120, len(qs)).astype(np.int32))

    ctx = bwg.Context(0)
    ctx.set_option(bwg.OPT_AUTOTUNE, 0)  # the default mapping: queries up to 32 columns one per lane
    try:
        for name, batch in (("accepted", build(60000, True)), ("declined", build(100000, False))):
            want = ob.extend_batch(batch, mat, 6, 1, 6, 1, 100, 5, 100, nthreads=16)
            st = ctx.stage(batch, params)
            for run in range(3):
                _say("  batch of %s short queries, run %d" % (name, run))
                st.run()
                got = st.results()
                assert _first_bad(got, want) is None, (name, run, _first_bad(got, want))
                if name == "declined" and run == 0:
                    # the scenario under test: the short bin ran on a grid sized from batch A
                    short = [r for r in st.report() if r["kernel"].startswith("ext_thread<")]
                    assert short and max(r["grid"] for r in short) < 100, short
            st.close()
    finally:
        ctx.close()
