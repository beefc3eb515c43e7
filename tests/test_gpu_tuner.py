"""The autotuner and every mapping it can choose: each candidate must give the oracle's results
bit for bit, so the tuner's choice changes speed only."""
import re
import sys

import numpy as np
import pytest

import bwg
import oracle_binding as ob

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


def _first_bad(got, want):
    bad = np.nonzero(got != want)[0]
    return None if bad.size == 0 else (int(bad[0]), got[bad[0]], want[bad[0]])


def test_forced_extension_mappings_agree(gpu_ctx):
    mat = bwg.default_mat()
    batch = bwg.generate(bwg.WORKLOAD_EXT_TAIL, 30000, seed=21)
    params = bwg.ext_params(5, mat, 6, 1, 6, 1, 100, 5, 100)
    want = ob.extend_batch(batch, mat, 6, 1, 6, 1, 100, 5, 100)
    gpu_ctx.set_form(bwg.FORM_BATCHED)
    gpu_ctx.set_option(bwg.OPT_AUTOTUNE, 0)
    try:
        # thread kernel up to q=32 / up to q=16 / never; group kernels at the minimal K or one
        # above, with 16 or 8 lanes per job; 16- and 32-bit column state
        for tmax, kstep, width, lanes in [(32, 0, 16, 16), (16, 0, 16, 16), (0, 0, 16, 16), (32, 1, 16, 16),
                                          (0, 1, 16, 16), (32, 0, 32, 16), (0, 1, 32, 16), (32, 0, 16, 8),
                                          (0, 0, 16, 8), (0, 1, 16, 8), (0, 0, 32, 8), (32, 1, 32, 8)]:
            _say("forced mapping thread_max_q=%d k_step=%d width=%d lanes=%d" % (tmax, kstep, width, lanes))
            gpu_ctx.set_option(bwg.OPT_EXT_THREAD_MAX_Q, tmax)
            gpu_ctx.set_option(bwg.OPT_EXT_K_STEP, kstep)
            gpu_ctx.set_option(bwg.OPT_EXT_SCORE_WIDTH, width)
            gpu_ctx.set_option(bwg.OPT_EXT_GROUP_LANES, lanes)
            gpu_ctx.set_option(bwg.OPT_AUTOTUNE, 0)
            got = gpu_ctx.extend_arrays(batch, params)
            assert _first_bad(got, want) is None, (tmax, kstep, width, lanes, _first_bad(got, want))
            if lanes == 8 and tmax == 0:
                st = gpu_ctx.stage(batch, params)
                st.run()
                assert any("ext_group8<" in r["kernel"] for r in st.report())
                st.close()
    finally:
        gpu_ctx.set_option(bwg.OPT_EXT_THREAD_MAX_Q, 32)
        gpu_ctx.set_option(bwg.OPT_EXT_K_STEP, 0)
        gpu_ctx.set_option(bwg.OPT_EXT_SCORE_WIDTH, 16)
        gpu_ctx.set_option(bwg.OPT_EXT_GROUP_LANES, 16)
        gpu_ctx.set_option(bwg.OPT_AUTOTUNE, 1)


def _long_query_batch(n, rng, qlo, qhi, h0_of):
    """Reads of qlen in [qlo, qhi) against targets drawn from them with 3% substitutions and short
    indels, followed by up to 80 random bases."""
    qs, ts, h0s = [], [], []
    for _ in range(n):
        ql = int(rng.integers(qlo, qhi))
        q = rng.integers(0, 4, ql, dtype=np.uint8)
        t = []
        k = 0
        while k < ql:
            u = rng.random()
            if u < 0.03:
                t.append(int(rng.integers(0, 4)))
                k += 1
            elif u < 0.035:
                t.extend(int(x) for x in rng.integers(0, 4, int(rng.integers(1, 6))))
            elif u < 0.04:
                k += int(rng.integers(1, 6))
            else:
                t.append(int(q[k]))
                k += 1
        t.extend(int(x) for x in rng.integers(0, 4, int(rng.integers(0, 80))))
        qs.append(q)
        ts.append(np.asarray(t, dtype=np.uint8))
        h0s.append(h0_of(ql))
    return bwg.JobBatch.from_pairs(qs, ts, np.asarray(h0s, dtype=np.int32))


@pytest.mark.parametrize("width,lanes", [(16, 16), (32, 16), (16, 8), (32, 8)])
def test_long_queries_and_state_width_boundary(gpu_ctx, width, lanes):
    """Queries up to 500 columns in the group kernels (windows up to 33 columns per lane at w=250),
    with h0 + qlen*max_sc around 65536, where 16-bit column state stops being exact and the
    planner switches those jobs to 32-bit state."""
    rng = np.random.default_rng(91)
    batches = [
        (_long_query_batch(200, rng, 257, 500, lambda ql: int(rng.integers(1, 200))), 100),
        (_long_query_batch(200, rng, 257, 500, lambda ql: int(rng.integers(1, 200))), 250),
        (_long_query_batch(200, rng, 100, 500, lambda ql: int(65535 - ql + rng.integers(-3, 4))), 250),
    ]
    mat = bwg.default_mat()
    gpu_ctx.set_form(bwg.FORM_BATCHED)
    gpu_ctx.set_option(bwg.OPT_AUTOTUNE, 0)
    try:
        gpu_ctx.set_option(bwg.OPT_EXT_SCORE_WIDTH, width)
        gpu_ctx.set_option(bwg.OPT_EXT_GROUP_LANES, lanes)
        for bi, (batch, w) in enumerate(batches):
            _say("long queries batch %d w=%d width=%d lanes=%d" % (bi, w, width, lanes))
            got = gpu_ctx.extend_arrays(batch, bwg.ext_params(5, mat, 6, 1, 6, 1, w, 5, 100))
            want = ob.extend_batch(batch, mat, 6, 1, 6, 1, w, 5, 100)
            assert _first_bad(got, want) is None, (bi, _first_bad(got, want))
    finally:
        gpu_ctx.set_option(bwg.OPT_EXT_SCORE_WIDTH, 16)
        gpu_ctx.set_option(bwg.OPT_EXT_GROUP_LANES, 16)
        gpu_ctx.set_option(bwg.OPT_AUTOTUNE, 1)


def test_forced_rescue_mappings_agree(gpu_ctx):
    mat = bwg.default_mat()
    batch = bwg.generate(bwg.WORKLOAD_RESCUE, 3000, seed=22)
    params = bwg.ksw_params(5, mat, 6, 1, 6, 1, 20, 1)
    want = ob.align2_batch(batch, mat, 6, 1, 6, 1, 20, 1)
    gpu_ctx.set_form(bwg.FORM_BATCHED)
    gpu_ctx.set_option(bwg.OPT_AUTOTUNE, 0)
    try:
        # minimal K or one above; packed 16-bit pairs where exact (16 or 8 lanes per job), or 32-bit
        # state always
        for kstep, width, lanes in [(0, 16, 16), (1, 16, 16), (0, 32, 16), (1, 32, 16), (0, 16, 8), (1, 16, 8)]:
            _say("forced rescue k_step=%d width=%d lanes=%d" % (kstep, width, lanes))
            gpu_ctx.set_option(bwg.OPT_KSW_K_STEP, kstep)
            gpu_ctx.set_option(bwg.OPT_KSW_SCORE_WIDTH, width)
            gpu_ctx.set_option(bwg.OPT_KSW_GROUP_LANES, lanes)
            gpu_ctx.set_option(bwg.OPT_AUTOTUNE, 0)
            got = gpu_ctx.align2_arrays(batch, params)
            assert _first_bad(got, want) is None, (kstep, width, lanes, _first_bad(got, want))
            if lanes == 8:
                st = gpu_ctx.stage(batch, params)
                st.run()
                assert any(r["kernel"].startswith("ksw_pk8<") for r in st.report())
                st.close()
    finally:
        gpu_ctx.set_option(bwg.OPT_KSW_K_STEP, 0)
        gpu_ctx.set_option(bwg.OPT_KSW_SCORE_WIDTH, 16)
        gpu_ctx.set_option(bwg.OPT_KSW_GROUP_LANES, 16)
        gpu_ctx.set_option(bwg.OPT_AUTOTUNE, 1)


def _rescue_near_bound(rng, n, lo, hi):
    """Queries of lo..hi bases placed (mutated) inside random windows, so the best local score
    reaches qlen * match for the exact copies: the packed kernel's 16-bit bound is met from both
    sides."""
    qs, ts = [], []
    for k in range(n):
        L = int(rng.integers(lo, hi + 1))
        q = rng.integers(0, 4, L).astype(np.uint8)
        body = q.copy()
        if k % 3:
            for _ in range(int(rng.integers(1, 6))):
                pos = int(rng.integers(0, len(body)))
                op = int(rng.integers(0, 3))
                if op == 0:
                    body[pos] = (body[pos] + 1) % 4
                elif op == 1:
                    body = np.delete(body, pos)
                else:
                    body = np.insert(body, pos, rng.integers(0, 4, int(rng.integers(1, 4))).astype(np.uint8))
        if k % 17 == 0:
            q[int(rng.integers(0, L))] = 4  # an N
        pre = rng.integers(0, 4, int(rng.integers(0, 80))).astype(np.uint8)
        post = rng.integers(0, 4, int(rng.integers(0, 80))).astype(np.uint8)
        qs.append(q)
        ts.append(np.concatenate([pre, body, post]))
    return bwg.JobBatch.from_pairs(qs, ts)


@pytest.mark.parametrize("a,b,o,e,lo,hi", [(1, 4, 6, 1, 100, 256), (2, 3, 5, 2, 200, 256),
                                           (3, 4, 6, 1, 130, 200), (4, 2, 0, 16383, 100, 150),
                                           (1, 4, 16000, 300, 100, 256), (2, 3, 16383, 1, 200, 256)])
def test_rescue_score_width_boundary(gpu_ctx, a, b, o, e, lo, hi):
    """Jobs on both sides of the 16-bit bound (qlen * match < 500) in one batch, with ordinary and
    extreme gap penalties (the last set is one past the packed kernel's gap limit): both widths
    agree with the oracle."""
    _say("rescue width boundary a=%d b=%d o=%d e=%d" % (a, b, o, e))
    rng = np.random.default_rng(1000 + a * 7 + o)
    batch = _rescue_near_bound(rng, 1500, lo, hi)
    mat = bwg.default_mat(a, b)
    params = bwg.ksw_params(5, mat, o, e, o, e, 20, a)
    want = ob.align2_batch(batch, mat, o, e, o, e, 20, a)
    gpu_ctx.set_form(bwg.FORM_BATCHED)
    gpu_ctx.set_option(bwg.OPT_AUTOTUNE, 0)
    try:
        # the 8-lane packed form applies only with e < 1000; otherwise the option changes nothing
        for width, lanes in ((16, 16), (32, 16), (16, 8)):
            gpu_ctx.set_option(bwg.OPT_KSW_SCORE_WIDTH, width)
            gpu_ctx.set_option(bwg.OPT_KSW_GROUP_LANES, lanes)
            gpu_ctx.set_option(bwg.OPT_AUTOTUNE, 0)
            got = gpu_ctx.align2_arrays(batch, params)
            assert _first_bad(got, want) is None, (width, lanes, _first_bad(got, want))
    finally:
        gpu_ctx.set_option(bwg.OPT_KSW_SCORE_WIDTH, 16)
        gpu_ctx.set_option(bwg.OPT_KSW_GROUP_LANES, 16)
        gpu_ctx.set_option(bwg.OPT_AUTOTUNE, 1)


def test_autotuner_tunes_once_and_stays_exact():
    _say("autotuner table")
    ctx = bwg.Context(0)
    try:
        mat = bwg.default_mat()
        ext = bwg.generate(bwg.WORKLOAD_EXT_TAIL, 200000, seed=23)
        ep = bwg.ext_params(5, mat, 6, 1, 6, 1, 100, 5, 100)
        resc = bwg.generate(bwg.WORKLOAD_RESCUE, 5000, seed=23)
        kp = bwg.ksw_params(5, mat, 6, 1, 6, 1, 20, 1)
        assert ctx.tune_report() == ""
        ext_want = ob.extend_batch(ext, mat, 6, 1, 6, 1, 100, 5, 100)
        ext_got = ctx.extend_arrays(ext, ep)
        assert _first_bad(ext_got, ext_want) is None
        got = ctx.align2_arrays(resc, kp)
        assert _first_bad(got, ob.align2_batch(resc, mat, 6, 1, 6, 1, 20, 1)) is None
        rep = ctx.tune_report()
        _say(rep)
        lines = [l for l in rep.splitlines() if l]
        assert any(l.startswith("extend qlen") for l in lines)   # short queries: lanes per job
        assert any(l.startswith("extend window") for l in lines)  # group kernels: columns per lane
        assert any(l.startswith("rescue") for l in lines)
        for l in lines:
            assert l.count("*") == 1, l
            times = [float(x) for x in re.findall(r"([0-9.]+) ns/job", l)]
            assert len(times) >= 2 and all(t > 0 for t in times), l
            chosen = float(re.search(r"\*\S+ ([0-9.]+) ns/job", l).group(1))
            if l.startswith("extend"):
                # extension classes: the first candidate within 3% of the fastest (10% for the
                # short-query classes); the report rounds to 0.001 ns
                tie = 1.10 if l.startswith("extend qlen") else 1.03
                k = times.index(chosen)
                eps = 0.0015
                assert chosen <= tie * min(times) + eps, l
                assert all(t > tie * min(times) - eps for t in times[:k]), l
            else:
                assert chosen == min(times), l
        # a second batch with the same scoring reuses the table and gives the same results
        again = ctx.extend_arrays(ext, ep)
        assert ctx.tune_report() == rep
        assert np.array_equal(again, ext_got)
    finally:
        ctx.close()
