"""Pins the oracle's banded global alignment with traceback (orc_ksw_global2, oracle/ksw_global.c),
the checker of the GPU CIGAR DP.

Checks: the reference's two unit tests (sw.rs:1358-1379); the score against exhaustive enumeration
of global alignments under bwa's gap rules whenever the band covers both sequences; and the CIGAR
itself: it consumes both sequences exactly and re-scores to the returned score.  A wrong tie
order in the traceback bits, a dropped continuation bit or an off-by-one in the band-relative
storage breaks the re-scoring or the consumed lengths.
"""
import numpy as np
import pytest

import oracle_binding as ob

M, I, D = 0, 1, 2


def _runs(cig):
    return [(c & 0xf, c >> 4) for c in cig]


def _rescore(q, t, cig, mat, o_del, e_del, o_ins, e_ins):
    i = j = 0
    s = 0
    for op, n in _runs(cig):
        if op == M:
            for _ in range(n):
                s += int(mat[t[i] * 5 + q[j]])
                i += 1
                j += 1
        elif op == I:
            s -= o_ins + e_ins * n
            j += n
        else:
            s -= o_del + e_del * n
            i += n
    return s, i, j


def test_kat_exact_match_is_all_m():  # crates/bwa-extend/src/sw.rs:1358-1365
    s = [0, 1, 2, 3, 0, 1, 2, 3, 0, 1]
    score, cig = ob.global2(s, s, ob.scmat(1, 4), 6, 1, 6, 1, 100)
    assert score == len(s) and cig == [len(s) << 4]


def test_kat_single_deletion():  # sw.rs:1367-1379
    q = [0, 1, 2, 3, 0, 1, 2, 3]
    t = q[:4] + [2] + q[4:]
    _, cig = ob.global2(q, t, ob.scmat(1, 4), 6, 1, 6, 1, 100)
    assert sum(n for op, n in _runs(cig) if op == D) == 1
    assert sum(n for op, n in _runs(cig) if op == M) == len(q)


@pytest.mark.parametrize("a,b,o_del,e_del,o_ins,e_ins", [(1, 4, 6, 1, 6, 1), (2, 3, 5, 2, 3, 1),
                                                         (1, 1, 0, 1, 4, 3), (3, 2, 2, 1, 7, 1)])
def test_score_equals_exhaustive_enumeration(a, b, o_del, e_del, o_ins, e_ins):
    rng = np.random.default_rng(a * 100 + b * 10 + o_ins)
    mat = ob.scmat(a, b)
    for _ in range(300):
        ql, tl = int(rng.integers(1, 7)), int(rng.integers(1, 7))
        q = rng.integers(0, 5, ql).astype(np.uint8)
        t = rng.integers(0, 5, tl).astype(np.uint8)
        score, cig = ob.global2(q, t, mat, o_del, e_del, o_ins, e_ins, max(ql, tl))
        assert score == ob.brute_global(q, t, mat, o_del, e_del, o_ins, e_ins), (q, t)
        assert _rescore(q, t, cig, mat, o_del, e_del, o_ins, e_ins) == (score, tl, ql), (q, t, cig)


@pytest.mark.parametrize("w", [0, 1, 3, 10, 40])
def test_banded_cigar_consumes_both_and_rescores(w):
    rng = np.random.default_rng(w + 7)
    mat = ob.scmat(1, 4)
    for _ in range(200):
        ql = int(rng.integers(1, 120))
        q = rng.integers(0, 4, ql).astype(np.uint8)
        t = q.copy()
        for _ in range(int(rng.integers(0, 4))):  # a few indels and substitutions
            p = int(rng.integers(0, len(t)))
            r = int(rng.integers(0, 3))
            if r == 0:
                t[p] = (t[p] + 1) % 4
            elif r == 1 and len(t) > 1:
                t = np.delete(t, p)
            else:
                t = np.insert(t, p, int(rng.integers(0, 4)))
        tl = len(t)
        score, cig = ob.global2(q, t, mat, 6, 1, 6, 1, w)
        # with a path inside the band and the traceback starting at the last cell, the CIGAR is a
        # full global alignment of that score (a band too narrow leaves an unreachable score)
        if ql <= tl + w and abs(tl - ql) <= w:
            assert score > -(1 << 29)
            assert _rescore(q, t, cig, mat, 6, 1, 6, 1) == (score, tl, ql), (w, q, t, cig)
        runs = _runs(cig)
        assert all(n > 0 for _, n in runs)
        assert all(runs[k][0] != runs[k + 1][0] for k in range(len(runs) - 1)), cig


def test_empty_sides():
    mat = ob.scmat(1, 4)
    assert ob.global2([], [0, 1, 2], mat, 6, 1, 6, 1, 5) == (-(6 + 3), [(3 << 4) | D])
    assert ob.global2([0, 1], [], mat, 6, 1, 6, 1, 5) == (-(6 + 2), [(2 << 4) | I])
