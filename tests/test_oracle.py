"""Pins the CPU oracle (the checker every GPU parity test relies on).

Each check is chosen so that a plausible restatement mistake fails it: the reference's own
known-answer tests, the reference's brute-force comparison, closed forms, exhaustive comparisons,
hand-worked tracker sequences, and agreement with the independent lockstep restatement on the
reference's own acceptance-gate sweeps.
"""
import numpy as np
import pytest

import bwg
import lcg_gates as G
import oracle_binding as ob

INT_MAX = 2**31 - 1


# ---- known answers from the reference's unit tests ----------------------------------------------

def test_kat_align2_perfect_match():  # crates/bwa-extend/src/sw.rs:1214-1223
    q = [0, 1, 2, 3, 0, 1, 2, 3]
    r = ob.align2(q, q, ob.scmat(1, 4), 6, 1, 6, 1, 4, 1, 16)
    assert r[0] == 8 and (r[1], r[2]) == (0, 7) and (r[3], r[4]) == (0, 7)


def test_kat_align2_local_trims_flanks():  # sw.rs:1225-1237
    r = ob.align2([3, 3, 0, 0, 1, 1, 3, 3], [2, 0, 0, 1, 1, 2], ob.scmat(1, 4), 6, 1, 6, 1, 1, 1, 16)
    assert r[0] == 4 and (r[1], r[2]) == (2, 5) and (r[3], r[4]) == (1, 4)


def test_kat_exact_match_full_length():  # sw.rs:1313-1323
    s = [0, 1, 2, 3] * 4
    r = ob.extend2(s, s, ob.scmat(1, 4), 6, 1, 6, 1, 100, 0, 100, 1)
    assert (r[0], r[1], r[2], r[4]) == (17, 16, 16, 17)


def test_kat_zdrop_stops_runaway():  # sw.rs:1382-1393
    q = [0] * 10 + [1] * 40
    t = [0] * 10 + [2] * 40
    r = ob.extend2(q, t, ob.scmat(1, 4), 6, 1, 6, 1, 100, 0, 100, 1)
    assert r[0] == 11 and r[2] == 10


def test_kat_padded_qlen():  # crates/bwa-neon/src/matesw.rs:5168
    L = ob.lib()
    assert L.orc_padded_qlen(147, 1) == 160
    assert L.orc_lanes_for(249, 1) == 16 and L.orc_lanes_for(250, 1) == 8
    assert L.orc_padded_qlen(150, 1) == 160 and L.orc_padded_qlen(250, 1) == 256
    assert L.orc_padded_qlen(30, 10) == 32  # 300 >= 250: 8 lanes


# ---- brute force -------------------------------------------------------------------------------

def test_matches_unbanded_reference():  # sw.rs:1325-1356, same xorshift stream
    mat = ob.scmat(1, 4)
    n = 0
    for q, t in G.unbanded_cases():
        got = ob.extend2(q, t, mat, 6, 1, 6, 1, 1000, 0, 1_000_000, 1)[0]
        assert got == ob.brute_extend_unbanded(q, t, mat, 6, 1, 6, 1, 1)
        n += 1
    assert n == 500


def test_extend_unbanded_brute_scoring_sweep():
    # Positive-prefix targets keep every row alive, so band/z-drop/zero-row exits cannot fire and
    # the banded kernel must equal the unbanded recurrence for every scoring.
    rng = np.random.default_rng(7)
    for _ in range(300):
        a, b = int(rng.integers(1, 5)), int(rng.integers(1, 7))
        od, ed, oi, ei = (int(x) for x in rng.integers(1, 9, 2).tolist() + rng.integers(1, 4, 2).tolist())
        od, oi, ed, ei = od, oi, ed, ei
        n = int(rng.integers(20, 60))
        q = rng.integers(0, 4, n).tolist()
        t = list(q)
        if n > 25:
            p = n - 3 - int(rng.integers(0, 3))
            t[p] = (t[p] + 1) % 4
        h0 = 60
        mat = ob.scmat(a, b)
        got = ob.extend2(q, t, mat, od, ed, oi, ei, 10_000, 0, 0, h0)[0]
        assert got == ob.brute_extend_unbanded(q, t, mat, od, ed, oi, ei, h0)


def test_local_fwd_equals_gotoh_with_padding():
    rng = np.random.default_rng(11)
    for _ in range(400):
        a, b = int(rng.integers(1, 4)), int(rng.integers(1, 6))
        od, ed, oi, ei = int(rng.integers(1, 9)), int(rng.integers(1, 4)), int(rng.integers(1, 9)), int(rng.integers(1, 4))
        ql, tl = int(rng.integers(1, 41)), int(rng.integers(1, 81))
        q = rng.integers(0, 5, ql).tolist()
        t = rng.integers(0, 5, tl).tolist()
        if rng.random() < 0.5 and tl >= ql:  # plant the query
            at = int(rng.integers(0, tl - ql + 1))
            t[at:at + ql] = q
        mat = ob.scmat(a, b)
        lanes = ob.lib().orc_lanes_for(ql, a)
        pad = ob.lib().orc_padded_qlen(ql, a)
        s = ob.local_fwd(q, t, mat, od, ed, oi, ei, INT_MAX, INT_MAX, a, lanes)[0]
        assert s == ob.brute_gotoh_local(q, t, mat, od, ed, oi, ei, pad_to=pad)


def test_local_fwd_equals_path_enumeration():
    # Every local path enumerated explicitly on tiny inputs: score, first row reaching it (te) and
    # smallest column in that row (qe).  lanes=1 leaves the query unpadded.
    rng = np.random.default_rng(3)
    for _ in range(400):
        a, b = int(rng.integers(1, 4)), int(rng.integers(1, 6))
        od, ed, oi, ei = int(rng.integers(0, 5)), int(rng.integers(1, 3)), int(rng.integers(0, 5)), int(rng.integers(1, 3))
        q = rng.integers(0, 5, int(rng.integers(1, 7))).tolist()
        t = rng.integers(0, 5, int(rng.integers(1, 7))).tolist()
        mat = ob.scmat(a, b)
        s, te, qe, _, _ = ob.local_fwd(q, t, mat, od, ed, oi, ei, INT_MAX, INT_MAX, a, 1)
        assert (s, te, qe) == ob.brute_enum_local(q, t, mat, od, ed, oi, ei)


# ---- closed forms and exhaustive comparisons ----------------------------------------------------

def test_extend_perfect_diagonal_closed_form():  # batched.rs:643-655
    rng = np.random.default_rng(5)
    for _ in range(200):
        a, b = int(rng.integers(1, 5)), int(rng.integers(1, 7))
        n = int(rng.integers(1, 200))
        h0 = int(rng.integers(1, 100))
        q = rng.integers(0, 4, n).tolist()
        r = ob.extend2(q, q, ob.scmat(a, b), 6, 1, 6, 1, int(rng.integers(1, 150)),
                       int(rng.integers(0, 10)), int(rng.integers(0, 150)), h0)
        assert r == (h0 + n * a, n, n, n, h0 + n * a, 0)


def test_extend_empty_target():
    assert ob.extend2([0, 1, 2], [], ob.scmat(1, 4), 6, 1, 6, 1, 100, 5, 100, 30) == (30, 0, 0, 0, -1, 0)


def _clamp_f64(w, qlen, max_sc, eb, o_ins, e_ins, o_del, e_del):
    """The reference's f64 formula (sw.rs:296-308), vectorised: trunc toward zero, floor at 1."""
    qlen = np.asarray(qlen, dtype=np.float64)
    mi = ((qlen * max_sc + eb - o_ins) / e_ins + 1.0).astype(np.int64)
    md = ((qlen * max_sc + eb - o_del) / e_del + 1.0).astype(np.int64)
    return np.minimum(np.minimum(w, np.maximum(mi, 1)), np.maximum(md, 1))


def _clamp_int(w, qlen, max_sc, eb, o_ins, e_ins, o_del, e_del):
    qlen = np.asarray(qlen, dtype=np.int64)
    ni = qlen * max_sc + eb - o_ins
    nd = qlen * max_sc + eb - o_del
    a = np.where(ni <= 0, 1, ni // e_ins + 1)
    b = np.where(nd <= 0, 1, nd // e_del + 1)
    return np.minimum(np.minimum(w, a), b)


def test_band_clamp_integer_form_exhaustive():
    q = np.arange(0, 1 << 16)
    for max_sc in (-1, 0, 1, 2, 3, 5, 10, 40):
        for eb in (0, 5, 9, 100):
            for o_ins, o_del in ((1, 1), (6, 6), (8, 3), (40, 17)):
                for e_ins, e_del in ((1, 1), (2, 3), (3, 1), (7, 13)):
                    for w in (1, 30, 100, 1 << 30):
                        f = _clamp_f64(w, q, max_sc, eb, o_ins, e_ins, o_del, e_del)
                        i = _clamp_int(w, q, max_sc, eb, o_ins, e_ins, o_del, e_del)
                        assert np.array_equal(f, i), (max_sc, eb, o_ins, o_del, e_ins, e_del, w)
    L = ob.lib()
    rng = np.random.default_rng(1)
    for _ in range(2000):
        args = [int(rng.integers(1, 300)), int(rng.integers(0, 5000)), int(rng.integers(-2, 12)),
                int(rng.integers(0, 20)), int(rng.integers(0, 30)), int(rng.integers(1, 9)),
                int(rng.integers(0, 30)), int(rng.integers(1, 9))]
        want = int(_clamp_f64(*args))
        assert L.orc_clamp_band(*args) == want
        assert L.orc_clamp_band_int(*args) == want


# ---- the suboptimal tracker, hand-worked ---------------------------------------------------------

def _tracker(rows, minsc):
    t = ob.Tracker()
    for i, v in enumerate(rows):
        t.push_row(i, v, minsc)
    return t


def test_tracker_merge_only_on_strict_improvement():
    t = _tracker([12, 15, 15, 9, 20, 25, 11], 10)
    assert t.entries() == [(15, 1), (25, 5)]
    assert t.finish(25, 5, 1) == (-1, -1)      # window [-20, 30] covers both
    assert t.finish(25, 5, 10) == (15, 1)      # window [2, 8]


def test_tracker_row_not_advanced_on_tie():
    t = _tracker([30, 30, 30, 30], 10)
    assert t.entries() == [(30, 0), (30, 2)]  # advancing on ties would give [(30, 3)]
    assert t.finish(30, 0, 15) == (-1, -1)     # window [-2, 2]


def test_tracker_first_strictly_greater_wins_and_sentinels():
    rows = [20] + [3] * 9 + [20]
    t = _tracker(rows, 5)
    assert t.entries() == [(20, 0), (20, 10)]
    assert t.finish(50, 30, 10) == (20, 0)
    assert t.finish(0, -1, 1) == (-1, -1)      # gmax == 0
    assert _tracker([10], 10).entries() == [(10, 0)]  # minsc itself is kept
    assert _tracker([9], 10).entries() == []


# ---- agreement with the independent lockstep restatement ----------------------------------------

def _invariants(r, h0, wclamp):
    score, qle, tle, gtle, gscore, max_off = r
    assert score >= h0
    assert (gscore == -1) == (gtle == 0)
    assert (qle == 0) == (tle == 0) == (score == h0)
    assert 0 <= max_off <= wclamp


def test_extend_single_gate_lockstep_and_invariants():
    L = ob.lib()
    for c in G.ext_single_cases():
        want = ob.extend2(c["q"], c["t"], c["mat"], c["o_del"], c["e_del"], c["o_ins"], c["e_ins"],
                          c["w"], c["end_bonus"], c["zdrop"], c["h0"])
        got = ob.lockstep_extend([c["q"]], [c["t"]], [c["h0"]], c["mat"], c["o_del"], c["e_del"],
                                 c["o_ins"], c["e_ins"], c["w"], c["end_bonus"], c["zdrop"])[0]
        assert got == want
        wc = L.orc_clamp_band(c["w"], len(c["q"]), max(c["mat"]), c["end_bonus"], c["o_ins"],
                              c["e_ins"], c["o_del"], c["e_del"])
        _invariants(want, c["h0"], wc)


@pytest.mark.parametrize("gen", [G.ext_batch_rounds, G.ext_read_sized, G.ext_ungapped])
def test_extend_batch_gates_lockstep(gen):
    for c in gen():
        args = (c["mat"], c["o_del"], c["e_del"], c["o_ins"], c["e_ins"], c["w"], c["end_bonus"], c["zdrop"])
        got = ob.lockstep_extend(c["queries"], c["targets"], c["h0s"], *args, lanes=8)
        for k in range(len(c["queries"])):
            assert got[k] == ob.extend2(c["queries"][k], c["targets"][k], *args, c["h0s"][k])


@pytest.mark.parametrize("a,b,minsc", G.RESCUE_SCORINGS)
def test_rescue_gate_lockstep_and_semantics(a, b, minsc):
    qs, ts = G.rescue_jobs()
    mat = ob.scmat(a, b)
    lock = ob.lockstep_local_fwd(qs, ts, [minsc] * len(qs), [INT_MAX] * len(qs), mat, 6, 1, 6, 1, a)
    populated = 0
    for k in range(len(qs)):
        lanes = ob.lib().orc_lanes_for(len(qs[k]), a)
        fwd = ob.local_fwd(qs[k], ts[k], mat, 6, 1, 6, 1, minsc, INT_MAX, a, lanes)
        assert lock[k] == fwd
        scal = ob.align2(qs[k], ts[k], mat, 6, 1, 6, 1, minsc, a, lanes)
        bat = ob.batched_align2_job(qs[k], ts[k], mat, 6, 1, 6, 1, minsc, a)
        assert scal == bat  # sw.rs ksw_align2 == matesw.rs batched_ksw_align2 per job
        assert scal[:1] + scal[2:3] + scal[4:] == (fwd[0], fwd[2], fwd[1], fwd[3], fwd[4])
        populated += fwd[3] >= 0
    assert populated > 50


def test_rescue_endsc_lockstep():
    qs, ts = G.rescue_jobs()
    qs, ts = qs[:400], ts[:400]
    mat = ob.scmat(1, 4)
    lock = ob.lockstep_local_fwd(qs, ts, [19] * 400, [60] * 400, mat, 6, 1, 6, 1, 1)
    for k in range(400):
        lanes = ob.lib().orc_lanes_for(len(qs[k]), 1)
        assert lock[k] == ob.local_fwd(qs[k], ts[k], mat, 6, 1, 6, 1, 19, 60, 1, lanes)


def test_rescue_invariants():
    qs, ts = G.rescue_jobs()
    mat = ob.scmat(1, 4)
    for k in range(0, 2000, 3):
        s, qb, qe, tb, te, s2, te2 = ob.batched_align2_job(qs[k], ts[k], mat, 6, 1, 6, 1, 19, 1)
        assert s >= 0 and s2 <= s
        assert (te == -1) == (s == 0)
        if s == 0:
            assert qe == -1
        if s < 19:
            assert qb == -1 and tb == -1
        if qb >= 0:
            assert 0 <= qb <= qe < len(qs[k]) and 0 <= tb <= te < len(ts[k])
        assert (s2 == -1) == (te2 == -1)


# ---- ungapped-diagonal fast path (crates/bwa-neon/src/batched.rs:602-697) -----------------------

def test_ungapped_threshold_matches_reference_cases():
    """x_threshold (batched.rs:756-764): 1 at bwa's defaults; the -B 3 and -B 2 -E 3 cases the
    reference documents (a strict win over the cheapest gap)."""
    assert ob.ungapped_xmax(ob.scmat(1, 4), 6, 1, 6, 1) == 1
    assert ob.ungapped_xmax(ob.scmat(1, 3), 6, 1, 6, 1) == 1   # 2*(1+3) = 8 > 6+1: X = 1, not 2
    assert ob.ungapped_xmax(ob.scmat(1, 2), 6, 3, 6, 3) == 2   # 3*(1+2) == 6+3 is a tie: X = 2
    assert ob.ungapped_xmax([0] * 25, 6, 1, 6, 1) == -1        # degenerate matrix disables it


def _assert_ungapped_exact(batch, mat, o_del, e_del, o_ins, e_ins, w, eb, zd):
    fast, hit = ob.ungapped_batch(batch, mat, o_del, e_del, o_ins, e_ins, zd)
    full = ob.extend_batch(batch, mat, o_del, e_del, o_ins, e_ins, w, eb, zd)
    bad = np.nonzero(hit & (fast != full))[0]
    assert bad.size == 0, "job %d: fast %s dp %s" % (bad[0], fast[bad[0]], full[bad[0]])
    return hit


def test_ungapped_accepts_only_dp_identical_results():
    """Every job the fast path accepts has exactly ksw_extend2's result (the reference's claim),
    on the synthetic workloads and on the gate sweeps with their varied scorings."""
    mat = bwg.default_mat()
    for kind in (bwg.WORKLOAD_EXT_SMALL, bwg.WORKLOAD_EXT_TAIL):
        b = bwg.generate(kind, 20000, seed=3)
        hit = _assert_ungapped_exact(b, mat, 6, 1, 6, 1, 100, 5, 100)
        assert hit.mean() > 0.3  # clean reads: most extensions are ungapped
    n_hit = 0
    for gen in (G.ext_ungapped, G.ext_read_sized):
        for case in gen():
            b = bwg.JobBatch.from_pairs(case["queries"], case["targets"], case["h0s"])
            hit = _assert_ungapped_exact(b, case["mat"], case["o_del"], case["e_del"], case["o_ins"],
                                         case["e_ins"], case["w"], case["end_bonus"], case["zdrop"])
            n_hit += int(hit.sum())
    assert n_hit > 100


def test_ungapped_closed_form_and_declines():
    mat = ob.scmat(1, 4)
    q = [0, 1, 2, 3] * 10
    b = bwg.JobBatch.from_pairs([q, q, q, q[:5] + [4] + q[6:], q], [q + [1] * 7, q[:-1], q, q, q[:20] + [1] + q[21:]],
                                [7, 7, 0 + 1, 7, 7])
    fast, hit = ob.ungapped_batch(b, mat, 6, 1, 6, 1, 100)
    assert hit[0] and tuple(fast[0]) == (7 + 40, 40, 40, 40, 47, 0)  # mismatch-free closed form
    assert not hit[1]   # target shorter than the query
    assert hit[2] and fast[2]["score"] == 41
    assert not hit[3]   # an N on the diagonal
    # one mismatch at column 20: best is the end of the walk, 7 + 39 - 4 = 42 at qle 40
    assert hit[4] and tuple(fast[4]) == (42, 40, 40, 40, 42, 0)
    full = ob.extend_batch(b, mat, 6, 1, 6, 1, 100, 0, 100)
    assert np.array_equal(fast[hit], full[hit])
