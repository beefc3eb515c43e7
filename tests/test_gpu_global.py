"""Banded global alignment with traceback on the device (bwg_global_batch) against the oracle's
ksw_global2 (orc_global_batch): every score and every CIGAR run, bit for bit."""
import sys

import numpy as np
import pytest

import bwg
import oracle_binding as ob

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


def _check(ctx, batch, w, a=1, b=4, o_del=6, e_del=1, o_ins=6, e_ins=1):
    mat = bwg.default_mat(a, b)
    p = bwg.global_params(5, mat, o_del, e_del, o_ins, e_ins)
    got_s, got_n, got_c = ctx.global_arrays(batch, p, w)
    want_s, want_n, want_c = ob.global_batch(batch, w, mat, o_del, e_del, o_ins, e_ins)
    bad = np.nonzero(got_s != want_s)[0]
    assert bad.size == 0, ("score", int(bad[0]), int(got_s[bad[0]]), int(want_s[bad[0]]))
    bad = np.nonzero(got_n != want_n.astype(np.uint32))[0]
    assert bad.size == 0, ("runs", int(bad[0]), int(got_n[bad[0]]), int(want_n[bad[0]]))
    assert np.array_equal(got_c, want_c)
    return want_s, want_n


@pytest.mark.parametrize("seed", [1, 2])
def test_region_jobs_match_oracle(gpu_ctx, seed):
    _say("global region jobs seed=%d" % seed)
    batch, w = bwg.generate_global_jobs(6000, seed=seed)
    s, n = _check(gpu_ctx, batch, w)
    assert (n > 1).sum() > 500  # jobs with gaps in their CIGAR


# (31, 32): scores at both ends of the packed 6-bit range; (1, 40): a mismatch score outside it,
# which runs the kernel form with five score planes per query column
@pytest.mark.parametrize("a,b,o_del,e_del,o_ins,e_ins", [(2, 3, 5, 2, 3, 1), (1, 1, 0, 1, 4, 3), (3, 2, 2, 1, 7, 1),
                                                         (31, 32, 6, 1, 6, 1), (1, 40, 6, 1, 6, 1)])
def test_scoring_sweep_matches_oracle(gpu_ctx, a, b, o_del, e_del, o_ins, e_ins):
    _say("global scoring %s" % ((a, b, o_del, e_del, o_ins, e_ins),))
    batch, _ = bwg.generate_global_jobs(3000, seed=a * 10 + b, max_indel=30)
    rng = np.random.default_rng(a + b)
    diff = np.abs(batch.q_len.astype(np.int64) - batch.t_len)
    w = (diff + rng.integers(0, 60, len(batch))).astype(np.int32)  # from the tightest band up
    _check(gpu_ctx, batch, w, a, b, o_del, e_del, o_ins, e_ins)


def test_edge_shapes_match_oracle(gpu_ctx):
    # empty sides, one-base sequences, bands of exactly the length difference, very wide bands,
    # queries up to 512 bases (the widest group kernels) and just past it, N bases
    _say("global edge shapes")
    rng = np.random.default_rng(7)
    qs, ts, ws = [], [], []
    for k in range(1500):
        ql = int(rng.choice([0, 1, 2, 7, 33, 150, 300, 512, 513, 700]))
        tl = max(0, ql + int(rng.integers(-5, 6))) if k % 3 else int(rng.integers(0, 40))
        q = rng.integers(0, 5, ql).astype(np.uint8)
        t = rng.integers(0, 5, tl).astype(np.uint8)
        if k % 2 and ql and tl:
            t[:min(ql, tl)] = q[:min(ql, tl)]
        qs.append(q)
        ts.append(t)
        d = abs(ql - tl)
        ws.append([d, d + 1, d + 3, 1000, 100000][k % 5])
    batch = bwg.JobBatch.from_pairs(qs, ts)
    _check(gpu_ctx, batch, np.asarray(ws, np.int32))


def test_staged_replays_and_rejections(gpu_ctx):
    _say("global replays")
    batch, w = bwg.generate_global_jobs(4000, seed=9)
    mat = bwg.default_mat()
    want_s, want_n, want_c = ob.global_batch(batch, w, mat, 6, 1, 6, 1)
    sb = gpu_ctx.stage_global(batch, bwg.global_params(5, mat, 6, 1, 6, 1), w)
    try:
        for rep in range(3):
            sb.run()
            assert np.array_equal(sb.results(), want_s), rep
            n, c = sb.cigars()
            assert np.array_equal(n, want_n.astype(np.uint32)) and np.array_equal(c, want_c), rep
    finally:
        sb.close()
    p = bwg.global_params(5, mat, 6, 1, 6, 1)
    bad_w = w.copy()
    k = int(np.argmax(np.abs(batch.q_len.astype(np.int64) - batch.t_len)))
    bad_w[k] = abs(int(batch.q_len[k]) - int(batch.t_len[k])) - 1  # the band misses the last cell
    if bad_w[k] >= 0:
        with pytest.raises(bwg.BwgError):
            gpu_ctx.global_arrays(batch, p, bad_w)
    # a code area beyond 1 GiB for one long query
    huge = bwg.JobBatch.from_pairs([np.zeros(40000, np.uint8)], [np.zeros(40000, np.uint8)])
    with pytest.raises(bwg.BwgError) as e:
        gpu_ctx.global_arrays(huge, p, np.array([20000], np.int32))
    assert e.value.code == bwg.BWG_E2BIG


@pytest.mark.parametrize("seed", [3, 4])
def test_long_queries_run_one_job_per_thread(gpu_ctx, seed):
    """Queries past the group kernel's 512 columns (up to 3000 bases, bands of 20-700) mixed with
    short ones in one batch: the long ones go to the one-job-per-thread kernel, whose scores and
    CIGARs must equal the oracle's as the group kernels' do."""
    _say("global long queries seed=%d" % seed)
    rng = np.random.default_rng(seed)
    qs, ts, ws = [], [], []
    for k in range(160):
        ql = int(rng.integers(513, 3000)) if k % 2 else int(rng.integers(1, 400))
        q = rng.integers(0, 4, ql).astype(np.uint8)
        t = list(q)
        for _ in range(int(rng.integers(0, 12))):  # substitutions and indels
            pos = int(rng.integers(0, len(t)))
            r = rng.random()
            if r < 0.5:
                t[pos] = int(rng.integers(0, 4))
            elif r < 0.75:
                t[pos:pos] = [int(x) for x in rng.integers(0, 4, int(rng.integers(1, 30)))]
            elif len(t) > 40:
                del t[pos:pos + int(rng.integers(1, 30))]
        t = np.asarray(t, np.uint8)
        qs.append(q)
        ts.append(t)
        d = abs(len(q) - len(t))
        ws.append(d + int(rng.choice([0, 5, 20, 100, 700])))
    batch = bwg.JobBatch.from_pairs(qs, ts)
    s, n = _check(gpu_ctx, batch, np.asarray(ws, np.int32))
    assert (n[1::2] > 1).sum() > 40  # long jobs with gaps
    st = gpu_ctx.stage_global(batch, bwg.global_params(5, bwg.default_mat(), 6, 1, 6, 1), np.asarray(ws, np.int32))
    try:
        st.run()
        assert any(r["kernel"] == "glb_plain" for r in st.report())
    finally:
        st.close()
