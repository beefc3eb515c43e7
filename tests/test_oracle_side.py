"""Pins the oracle's band doubling around ksw_extend2 (oracle/ksw_side.c), the checker of the
band-retry entry point.

The reference has no known-answer test for the retry rule, so each check below is one that a
plausible restatement mistake fails: the shift form of the 3/4-band threshold at odd w (a float or
`3*w/4` threshold accepts where bwa retries), the w = 1 degenerate case where only the score test
can accept, the score test itself (a right-side job that gains nothing), the last-try rule, and
agreement of the per-job form (extend_side) with the batched rounds (run_side) on a workload where
many jobs are retried.
"""
import numpy as np
import pytest

import bwg
import oracle_binding as ob

MAT = ob.scmat(1, 4)
SC = (6, 1, 6, 1)


def _gapped_pair(rng, qlen, cut, gap):
    """A query, and a target holding it with `gap` extra bases after position `cut`: the best
    extension runs past the deletion, so its best cell sits `gap` off the diagonal."""
    q = rng.integers(0, 4, qlen).astype(np.uint8)
    ins = ((q[cut:cut + gap] + 1) % 4).astype(np.uint8)  # never equal to the bases they displace
    t = np.concatenate([q[:cut], ins, q[cut:], rng.integers(0, 4, 20).astype(np.uint8)])
    return q, t


@pytest.mark.parametrize("gap,retried", [(4, True), (3, False)])
def test_threshold_is_the_shift_form_at_odd_w(gap, retried):
    # w = 7: (7 >> 1) + (7 >> 2) = 4, so max_off 4 is retried (0.75 * 7 = 5.25 would accept it)
    rng = np.random.default_rng(7 + gap)
    q, t = _gapped_pair(rng, 40, 10, gap)
    h0 = 20
    r7 = ob.extend2(q, t, MAT, *SC, 7, 5, 100, h0)
    assert r7[5] == gap and r7[0] == h0 + 40 - (6 + gap)  # 40 matches, one deletion of `gap`
    res, w = ob.extend_side(q, t, MAT, *SC, 7, 5, 100, h0, -1)
    if retried:
        assert w == 14 and res == ob.extend2(q, t, MAT, *SC, 14, 5, 100, h0)
    else:
        assert w == 7 and res == r7


def test_w1_accepts_only_on_score():
    # at w = 1 the offset test reads max_off < 0 and never fires: every job whose score differs from
    # prev is extended again at w = 2
    rng = np.random.default_rng(3)
    for _ in range(50):
        n = int(rng.integers(5, 40))
        q = rng.integers(0, 4, n).astype(np.uint8)
        t = np.concatenate([q, rng.integers(0, 4, 10).astype(np.uint8)])
        t[rng.integers(0, n)] ^= 1
        h0 = int(rng.integers(5, 60))
        res, w = ob.extend_side(q, t, MAT, *SC, 1, 5, 100, h0, -1)
        assert w == 2 and res == ob.extend2(q, t, MAT, *SC, 2, 5, 100, h0)


def test_score_equal_to_prev_accepts_at_first_band():
    # a right extension starting at the left score h0 that gains nothing (every base mismatches)
    # scores exactly h0 = prev and is accepted at the first band even at w = 1
    q = np.array([0, 1, 2, 3, 0, 1], np.uint8)
    t = np.array([2, 3, 0, 1, 2, 3, 0, 1], np.uint8)  # q + 2 (mod 4): all mismatches on the diagonal
    h0 = 3
    r = ob.extend2(q, t, MAT, *SC, 1, 5, 100, h0)
    assert r[0] == h0
    res, w = ob.extend_side(q, t, MAT, *SC, 1, 5, 100, h0, h0)
    assert w == 1 and res == r
    res, w = ob.extend_side(q, t, MAT, *SC, 1, 5, 100, h0, -1)  # a left job's prev never matches
    assert w == 2


def test_last_try_always_accepts():
    # from w = 1 (where only the score test could accept) a 2-base deletion is out of reach at w = 1,
    # reached at w = 2 with max_off 2 >= (2 >> 1) + (2 >> 2) = 1, and accepted at w = 4 by the
    # offset test (2 < 3): each try count stops where the rule says
    rng = np.random.default_rng(11)
    q, t = _gapped_pair(rng, 40, 10, 2)
    h0 = 30
    r1 = ob.extend2(q, t, MAT, *SC, 1, 5, 100, h0)
    r2 = ob.extend2(q, t, MAT, *SC, 2, 5, 100, h0)
    assert r2[0] != r1[0] and r2[5] == 2
    for tries, want_w in [(1, 1), (2, 2), (3, 4), (4, 4)]:
        res, w = ob.extend_side(q, t, MAT, *SC, 1, 5, 100, h0, -1, max_band_try=tries)
        assert w == want_w, (tries, w)
        assert res == ob.extend2(q, t, MAT, *SC, want_w, 5, 100, h0)


@pytest.mark.parametrize("w,tries", [(3, 2), (5, 2), (7, 3), (1, 2), (100, 2)])
def test_batched_rounds_equal_per_job_tries(w, tries):
    batch = bwg.generate(bwg.WORKLOAD_EXT_TAIL, 3000, seed=31 + w)
    prev = np.where(np.arange(len(batch)) % 2 == 0, -1, batch.h0).astype(np.int32)
    out, w_out = ob.run_side(batch, prev, MAT, *SC, w, 5, 100, max_band_try=tries)
    for k in range(0, len(batch), 7):
        q = batch.arena[batch.q_off[k]:batch.q_off[k] + batch.q_len[k]]
        t = batch.arena[batch.t_off[k]:batch.t_off[k] + batch.t_len[k]]
        res, wk = ob.extend_side(q, t, MAT, *SC, w, 5, 100, int(batch.h0[k]), int(prev[k]), tries)
        assert tuple(out[k]) == res and w_out[k] == wk, k
    assert set(np.unique(w_out)) <= {w << r for r in range(tries)}
    if w <= 5:
        assert (w_out != w).sum() > 20  # the sweep exercises the retry path
