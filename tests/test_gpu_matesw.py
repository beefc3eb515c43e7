"""Mate rescue with device-resident windows (bwg_matesw_batch) against the oracle's round
(orc_matesw_batch: matesw_collect, bns_fetch_seq and batched_ksw_align2 per job): every window and
every result field, bit for bit."""
import sys

import numpy as np
import pytest

import bwg
import oracle_binding as ob

pytestmark = pytest.mark.gpu

MAT = bwg.default_mat()


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


def _params(msl=19, a=1):
    return bwg.ksw_params(5, MAT, 6, 1, 6, 1, msl * a, a)


def _check(ctx, rnd):
    ctx.set_reference(rnd.fwd, rnd.contigs)
    got_w, got = ctx.matesw_arrays(rnd, _params(rnd.min_seed_len))
    want_w, want = ob.matesw_batch(rnd, MAT, 6, 1, 6, 1, rnd.min_seed_len, 1)
    bad = np.nonzero(got_w != want_w)[0]
    assert bad.size == 0, ("window", int(bad[0]), got_w[bad[0]], want_w[bad[0]])
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, ("result", int(bad[0]), got[bad[0]], want[bad[0]], want_w[bad[0]])
    return want_w, want


@pytest.mark.parametrize("seed", [3, 4])
def test_generated_round_matches_oracle(gpu_ctx, seed):
    _say("matesw generated round seed=%d" % seed)
    gpu_ctx.set_form(bwg.FORM_BATCHED)
    rnd = bwg.generate_rescue_round(20000, seed=seed)
    win, out = _check(gpu_ctx, rnd)
    assert ((win["flags"] & bwg.WIN_VALID) != 0).sum() > 15000
    assert (out["score"] >= 140).sum() > 5000  # the planted mates are found


def _edge_round(rng):
    L = 300_000
    fwd = rng.integers(0, 4, L).astype(np.uint8)
    fwd[rng.random(L) < 0.001] = 4
    starts = np.array([0, 1000, 150_000, 150_100, 299_000], np.int64)
    contigs = np.zeros(len(starts), dtype=bwg.CONTIG_DTYPE)
    contigs["offset"] = starts
    contigs["len"] = np.diff(np.concatenate([starts, [L]]))
    pes = np.zeros(4, dtype=bwg.PE_DTYPE)
    pes[0] = (50, 400, 0)
    pes[1] = (1, 900, 0)
    pes[2] = (100, 300, 0)
    pes[3] = (1, 2000, 1)
    pos = [0, 5, 999, 1000, 1001, 149_990, 150_050, 150_099, 298_990, 299_500, L - 1, L, L + 5, 2 * L - 1,
           2 * L - 200, 2 * L - 150_050]
    anchors, mates, ma = [], [], []
    for k in range(600):
        rb = pos[k % len(pos)] if k < 3 * len(pos) else int(rng.integers(0, 2 * L))
        fwd_rb = rb if rb < L else 2 * L - 1 - rb
        rid = int(np.searchsorted(starts, min(fwd_rb, L - 1), side="right") - 1)
        if k % 11 == 0:
            rid = (rid + 1) % len(starts)  # disagrees with every window
        l_ms = int(rng.choice([30, 100, 150, 151, 250, 300]))
        if rng.random() < 0.6 and rb < L - l_ms - 1000:
            x = rb + int(rng.integers(0, 800))
            m = fwd[x:x + l_ms][::-1].copy()
            m = np.where(m < 4, 3 - m, m).astype(np.uint8)  # a mate on the other strand downstream
        else:
            m = rng.integers(0, 5, l_ms).astype(np.uint8)
        begin = len(ma)
        for _ in range(int(rng.integers(0, 3))):
            ma.append(int(rng.integers(0, 2 * L)))
        anchors.append((rb, rid, begin, len(ma)))
        mates.append(m)
    arena = np.concatenate(mates)
    a = np.zeros(len(anchors), dtype=bwg.ANCHOR_DTYPE)
    o = 0
    for k, (rb, rid, b0, b1) in enumerate(anchors):
        a[k] = (rb, rid, len(mates[k]), o, b0, b1)
        o += len(mates[k])
    return bwg.RescueRound(fwd, contigs, pes, a, np.asarray(ma, np.int64), arena, 19)


def test_edge_round_matches_oracle(gpu_ctx):
    # genome ends, contig junctions and a 100-base contig, rid mismatches, N bases, mates longer
    # than the group kernels take (the plain bin), existing mate regions
    _say("matesw edge round")
    gpu_ctx.set_form(bwg.FORM_BATCHED)
    rnd = _edge_round(np.random.default_rng(9))
    win, out = _check(gpu_ctx, rnd)
    assert ((win["flags"] & bwg.WIN_VALID) != 0).sum() > 500
    for msl in (0, 400):
        rnd.min_seed_len = msl
        _check(gpu_ctx, rnd)


def test_replays_and_reference_guard(gpu_ctx):
    _say("matesw replays")
    gpu_ctx.set_form(bwg.FORM_BATCHED)
    rnd = bwg.generate_rescue_round(5000, seed=11)
    gpu_ctx.set_reference(rnd.fwd, rnd.contigs)
    want_w, want = ob.matesw_batch(rnd, MAT, 6, 1, 6, 1, 19, 1)
    sb = gpu_ctx.stage_matesw(rnd, _params())
    try:
        with pytest.raises(bwg.BwgError):  # the staged batch still reads the reference
            gpu_ctx.set_reference(rnd.fwd, rnd.contigs)
        for rep in range(3):
            sb.run()
            assert np.array_equal(sb.windows(), want_w), rep
            assert np.array_equal(sb.results(), want), rep
    finally:
        sb.close()
    gpu_ctx.set_reference(rnd.fwd, rnd.contigs)
    fresh = bwg.Context(0)
    try:
        with pytest.raises(bwg.BwgError):  # no reference on this context
            fresh.matesw_arrays(rnd, _params())
    finally:
        fresh.close()
