"""Pins the oracle's mate-rescue windows (oracle/ksw_matesw.c), the checker of bwg_matesw_batch.

The reference has no unit test for matesw_collect, so the checks are hand-worked cases from bwa's
definitions (pair orientations, the doubled coordinate space, contig clamps), brute-force
comparisons (pos2rid against a linear scan, the reverse half against the reverse complement) and a
planted mate whose rescue has a closed-form answer.  A swapped orientation bit, a missing strand
mirror or an off-by-one in the window would fail one of them.
"""
import numpy as np
import pytest

import bwg
import oracle_binding as ob

MAT = ob.scmat(1, 4)


def test_infer_dir_hand_worked():
    # l_pac = 1000; forward coordinates < 1000, reverse half 1000..1999
    assert ob.infer_dir(1000, 100, 600) == (0, 500)    # both forward, mate downstream: FF
    assert ob.infer_dir(1000, 100, 1400) == (1, 499)   # mate reverse at 2000-1-1400 = 599: FR
    assert ob.infer_dir(1000, 100, 50) == (3, 50)      # both forward, mate upstream: flipped FF
    assert ob.infer_dir(1000, 1400, 100) == (1, 499)   # anchor reverse: mate mirrored to 1899, downstream
    assert ob.infer_dir(1000, 1400, 1500) == (0, 100)


def _contigs(starts, lens):
    c = np.zeros(len(starts), dtype=bwg.CONTIG_DTYPE)
    c["offset"] = starts
    c["len"] = lens
    return c


def test_pos2rid_against_linear_scan():
    rng = np.random.default_rng(1)
    starts = np.array([0, 17, 18, 400, 1000, 2047], np.int64)
    c = _contigs(starts, np.diff(np.concatenate([starts, [3000]])))
    for p in list(range(0, 3000, 7)) + list(starts) + list(starts - 1)[1:] + [2999]:
        want = int(np.searchsorted(starts, p, side="right") - 1)
        assert ob.pos2rid(3000, c, int(p)) == want, p
    assert ob.pos2rid(3000, c, 3000) == -1 and ob.pos2rid(3000, c, 5000) == -1
    # a position before the first contig reports contig 0 (the last probe), as the reference does
    assert ob.pos2rid(3000, _contigs([10, 20], [10, 2980]), 3) == 0


def test_reverse_half_is_reverse_complement():
    rng = np.random.default_rng(2)
    fwd = rng.integers(0, 5, 257).astype(np.uint8)
    rev = ob.ref_bases(fwd, 257, 514)
    want = np.where(fwd[::-1] < 4, 3 - fwd[::-1], 4)
    assert np.array_equal(rev, want)
    assert np.array_equal(ob.ref_bases(fwd, 0, 257), fwd)


def test_fetch_window_clamps_to_the_midpoint_contig_both_strands():
    c = _contigs([0, 1000], [1000, 500])  # l_pac 1500, doubled space 3000
    # forward: a window over the junction clamps to the contig holding its midpoint
    assert ob.fetch_window(1500, c, 900, 1050, 1200) == (1000, 1200, 1)
    assert ob.fetch_window(1500, c, 900, 950, 1200) == (900, 1000, 0)
    # reverse half: contig 1 [1000, 1500) mirrors to [1500, 2000), contig 0 to [2000, 3000)
    assert ob.fetch_window(1500, c, 1900, 2050, 2200) == (2000, 2200, 0)
    assert ob.fetch_window(1500, c, 1900, 1950, 2200) == (1900, 2000, 1)
    # bounds given in either order
    assert ob.fetch_window(1500, c, 1200, 1050, 900) == (1000, 1200, 1)


def _round(fwd, contigs, pes, anchors, ma_rb, mates, msl=19):
    arena = np.concatenate(mates).astype(np.uint8)
    a = np.zeros(len(anchors), dtype=bwg.ANCHOR_DTYPE)
    o = 0
    for k, (rb, rid, begin, end) in enumerate(anchors):
        a[k] = (rb, rid, len(mates[k]), o, begin, end)
        o += len(mates[k])
    return bwg.RescueRound(fwd, contigs, pes, a, np.asarray(ma_rb, np.int64), arena, msl)


def test_orientation_windows_hand_worked():
    L = 100_000
    fwd = np.zeros(L, np.uint8)
    c = _contigs([0], [L])
    pes = np.zeros(4, dtype=bwg.PE_DTYPE)
    pes[:] = (100, 700, 0)
    mate = np.zeros(150, np.uint8)
    rnd = _round(fwd, c, pes, [(50_000, 0, 0, 0)], [], [mate])
    win, _ = ob.matesw_batch(rnd, MAT, 6, 1, 6, 1, 19, 1)
    # r = 0 FF larger: [rb+lo, rb+hi+l); r = 1 FR larger, mate reversed: [rb+lo-l, rb+hi);
    # r = 2 RF smaller, reversed: [rb-hi-l, rb-lo); r = 3 RR smaller: [rb-hi, rb-lo+l)
    want = [(50_100, 50_850, 0), (49_950, 50_700, 2), (49_150, 49_900, 2), (49_300, 50_050, 0)]
    for r, (rb, re, rev) in enumerate(want):
        assert (win[r]["rb"], win[r]["re"]) == (rb, re), r
        assert win[r]["flags"] == bwg.WIN_VALID | rev, r


def test_planted_mate_is_rescued_at_its_position():
    rng = np.random.default_rng(5)
    L = 200_000
    fwd = rng.integers(0, 4, L).astype(np.uint8)
    c = _contigs([0, 120_000], [120_000, 80_000])
    pes = np.zeros(4, dtype=bwg.PE_DTYPE)
    pes[:] = (1, 1000, 1)
    pes[1] = (1, 1000, 0)  # FR only
    x, isize, l = 30_000, 480, 150
    read2 = fwd[x + isize - l:x + isize][::-1]
    read2 = np.where(read2 < 4, 3 - read2, read2).astype(np.uint8)  # read 2 as sequenced
    read1 = fwd[x:x + l].copy()
    # anchor read 1 on the forward strand; anchor read 2 on the reverse half (mate read 1)
    rnd = _round(fwd, c, pes, [(x, 0, 0, 0), (2 * L - (x + isize), 0, 0, 0)], [], [read2, read1])
    win, out = ob.matesw_batch(rnd, MAT, 6, 1, 6, 1, 19, 1)
    for k, want_pos in [(0, x + isize - l), (1, 2 * L - (x + l))]:
        w, r = win[4 * k + 1], out[4 * k + 1]
        assert w["flags"] == bwg.WIN_VALID | bwg.WIN_REV
        assert r["score"] == l and r["qb"] == 0 and r["qe"] == l - 1
        assert w["rb"] + r["tb"] == want_pos and r["te"] - r["tb"] == l - 1
        for rr in (0, 2, 3):
            assert win[4 * k + rr]["flags"] & bwg.WIN_SKIPPED and tuple(out[4 * k + rr]) == (0, -1, -1, -1, -1, -1, -1)


def test_skip_rules():
    L = 100_000
    fwd = np.zeros(L, np.uint8)
    c = _contigs([0, 60_000], [60_000, 40_000])
    pes = np.zeros(4, dtype=bwg.PE_DTYPE)
    pes[:] = (100, 700, 0)
    mate = np.zeros(150, np.uint8)
    anchors = [(10_000, 0, 0, 1),          # mate region FR at distance 400: r = 1 skipped
               (10_000, 0, 1, 5),          # all four orientations paired: no call at all
               (59_500, 0, 5, 5),          # FF and FR windows run past the end of contig 0
               (10_000, 1, 5, 5),          # anchor rid disagrees with the window's contig
               (59_950, 0, 5, 5)]          # FF window entirely in contig 1
    ma = [2 * L - 1 - 10_400,
          10_300, 2 * L - 1 - 10_400, 2 * L - 1 - 9_600, 9_700]
    rnd = _round(fwd, c, pes, anchors, ma, [mate] * 5)
    win, _ = ob.matesw_batch(rnd, MAT, 6, 1, 6, 1, 19, 1)
    f = win["flags"].reshape(-1, 4)
    w = win.reshape(-1, 4)
    assert list(f[0] & bwg.WIN_SKIPPED) == [0, 4, 0, 0]
    assert all(f[1] & bwg.WIN_SKIPPED) and not any(f[1] & bwg.WIN_VALID)
    # midpoints 59_975 and 59_825 lie in contig 0: both windows end at its end
    assert (w[2][0]["rb"], w[2][0]["re"], w[2][0]["rid"]) == (59_600, 60_000, 0) and f[2][0] & bwg.WIN_VALID
    assert (w[2][1]["rb"], w[2][1]["re"], w[2][1]["rid"]) == (59_450, 60_000, 0) and f[2][1] & bwg.WIN_VALID
    assert not any(f[3] & bwg.WIN_VALID)
    assert w[4][0]["rid"] == 1 and not f[4][0] & bwg.WIN_VALID
    assert (w[4][0]["rb"], w[4][0]["re"]) == (60_050, 60_800)


@pytest.mark.parametrize("seed", [3, 4])
def test_generated_round_invariants(seed):
    rnd = bwg.generate_rescue_round(3000, seed=seed)
    win, out = ob.matesw_batch(rnd, MAT, 6, 1, 6, 1, 19, 1)
    valid = (win["flags"] & bwg.WIN_VALID) != 0
    assert valid.sum() > 2000
    assert np.all(win["rb"][valid] >= 0) and np.all(win["re"][valid] <= 2 * rnd.l_pac)
    assert np.all(win["re"][valid] - win["rb"][valid] >= rnd.min_seed_len)
    assert np.all(out["score"][~valid] == 0) and np.all(out["te"][~valid] == -1)
    # every valid job's result is batched_ksw_align2 on the fetched window
    for j in np.nonzero(valid)[0][::97]:
        k, r = divmod(int(j), 4)
        a = rnd.anchors[k]
        q = rnd.arena[a["ms_off"]:a["ms_off"] + a["l_ms"]]
        if win[j]["flags"] & bwg.WIN_REV:
            q = np.where(q[::-1] < 4, 3 - q[::-1], 4).astype(np.uint8)
        t = ob.ref_bases(rnd.fwd, int(win[j]["rb"]), int(win[j]["re"]))
        assert tuple(out[j]) == ob.batched_align2_job(q, t, MAT, 6, 1, 6, 1, 19, 1)
