"""The band-retry entry point (bwg_extend_side_batch) against the oracle's batched rounds
(orc_run_side, crates/bwa-mem/src/across.rs:1357-1460): every result field and every accepted band,
bit for bit."""
import sys

import numpy as np
import pytest

import bwg
import oracle_binding as ob

pytestmark = pytest.mark.gpu

MAT = bwg.default_mat()
SC = (6, 1, 6, 1)


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


def _first_bad(got, want):
    bad = np.nonzero(got != want)[0]
    return None if bad.size == 0 else (int(bad[0]), got[bad[0]], want[bad[0]])


def _check(ctx, batch, prev, w, tries, end_bonus=5, zdrop=100):
    params = bwg.ext_params(5, MAT, *SC, w, end_bonus, zdrop)
    got, got_w = ctx.extend_side_arrays(batch, params, prev, tries)
    want, want_w = ob.run_side(batch, prev, MAT, *SC, w, end_bonus, zdrop, max_band_try=tries)
    assert _first_bad(got, want) is None, (w, tries, _first_bad(got, want))
    assert np.array_equal(got_w, want_w), (w, tries, int(np.nonzero(got_w != want_w)[0][0]))
    return (want_w != w).sum()


def _prev(batch):
    # left jobs (-1) and right jobs (their h0, the left extension's score) alternate
    return np.where(np.arange(len(batch)) % 2 == 0, -1, batch.h0).astype(np.int32)


@pytest.mark.parametrize("w,tries", [(100, 2), (5, 2), (3, 2), (1, 2), (5, 1), (3, 3)])
def test_side_workload_matches_oracle(gpu_ctx, w, tries):
    _say("side batch long-tail w=%d tries=%d" % (w, tries))
    gpu_ctx.set_form(bwg.FORM_BATCHED)
    batch = bwg.generate(bwg.WORKLOAD_EXT_TAIL, 20000, seed=60 + w)
    retried = _check(gpu_ctx, batch, _prev(batch), w, tries)
    if w <= 5 and tries > 1:
        assert retried > 100


def _indel_heavy(rng, n, qmin, qmax, gmax):
    """Queries placed in targets with one deletion or insertion of up to gmax bases, so the best
    cell often sits far off the diagonal and the wider band finds a better score."""
    qs, ts, h0 = [], [], []
    for _ in range(n):
        L = int(rng.integers(qmin, qmax + 1))
        q = rng.integers(0, 4, L).astype(np.uint8)
        cut = int(rng.integers(1, L))
        g = int(rng.integers(0, gmax + 1))
        if rng.integers(0, 2):
            t = np.concatenate([q[:cut], rng.integers(0, 4, g).astype(np.uint8), q[cut:]])
        else:
            t = np.concatenate([q[:cut], q[min(L, cut + g):]])
        t = np.concatenate([t, rng.integers(0, 4, int(rng.integers(0, 100))).astype(np.uint8)])
        if len(t) == 0:
            t = np.array([0], np.uint8)
        qs.append(q)
        ts.append(t)
        h0.append(int(rng.integers(10, 120)))
    return bwg.JobBatch.from_pairs(qs, ts, h0)


@pytest.mark.parametrize("w,gmax,qmax", [(30, 40, 200), (100, 90, 250), (8, 12, 120)])
def test_side_indel_heavy_matches_oracle(gpu_ctx, w, gmax, qmax):
    _say("side batch indel-heavy w=%d gmax=%d" % (w, gmax))
    gpu_ctx.set_form(bwg.FORM_BATCHED)
    rng = np.random.default_rng(w * 101 + gmax)
    batch = _indel_heavy(rng, 6000, 33, qmax, gmax)
    retried = _check(gpu_ctx, batch, _prev(batch), w, 2)
    assert retried > 50


def test_side_plain_bin_jobs_retry(gpu_ctx):
    # queries longer than the group kernel's LDS capacity run in the plain bin, which reads the
    # active flags between rounds
    _say("side batch with plain-bin jobs")
    gpu_ctx.set_form(bwg.FORM_BATCHED)
    rng = np.random.default_rng(5)
    long_jobs = _indel_heavy(rng, 300, 300, 420, 12)
    short_jobs = _indel_heavy(rng, 3000, 33, 150, 12)
    qs = [long_jobs.arena[o:o + l] for o, l in zip(long_jobs.q_off, long_jobs.q_len)] + \
         [short_jobs.arena[o:o + l] for o, l in zip(short_jobs.q_off, short_jobs.q_len)]
    ts = [long_jobs.arena[o:o + l] for o, l in zip(long_jobs.t_off, long_jobs.t_len)] + \
         [short_jobs.arena[o:o + l] for o, l in zip(short_jobs.t_off, short_jobs.t_len)]
    h0 = np.concatenate([long_jobs.h0, short_jobs.h0])
    batch = bwg.JobBatch.from_pairs(qs, ts, h0)
    retried = _check(gpu_ctx, batch, _prev(batch), 6, 2)
    assert retried > 50


def test_side_graph_replays_and_serial_bins(gpu_ctx):
    _say("side batch replays")
    gpu_ctx.set_form(bwg.FORM_BATCHED)
    batch = bwg.generate(bwg.WORKLOAD_EXT_TAIL, 20000, seed=77)
    prev = _prev(batch)
    params = bwg.ext_params(5, MAT, *SC, 4, 5, 100)
    want, want_w = ob.run_side(batch, prev, MAT, *SC, 4, 5, 100)
    sb = gpu_ctx.stage_side(batch, params, prev, 2)
    try:
        for rep in range(4):  # the first run is eager, the later ones replay its captured graph
            sb.run()
            got = sb.results()
            assert _first_bad(got, want) is None, (rep, _first_bad(got, want))
            assert np.array_equal(sb.side_bands(), want_w), rep
    finally:
        sb.close()
    gpu_ctx.set_option(bwg.OPT_CONCURRENT_BINS, 0)
    try:
        got, got_w = gpu_ctx.extend_side_arrays(batch, params, prev, 2)
        assert _first_bad(got, want) is None and np.array_equal(got_w, want_w)
    finally:
        gpu_ctx.set_option(bwg.OPT_CONCURRENT_BINS, 1)


def test_side_rejects_bad_arguments(gpu_ctx):
    batch = bwg.generate(bwg.WORKLOAD_EXT_TAIL, 10, seed=1)
    prev = _prev(batch)
    for w, tries in [(100, 0), (100, 9), (-1, 2)]:
        params = bwg.ext_params(5, MAT, *SC, w, 5, 100)
        with pytest.raises(bwg.BwgError):
            gpu_ctx.extend_side_arrays(batch, params, prev, tries)
    params = bwg.ext_params(5, MAT, *SC, 1 << 30, 5, 100)
    with pytest.raises(bwg.BwgError):
        gpu_ctx.extend_side_arrays(batch, params, prev, 3)
