"""Pins the tie rules of the oracle's ksw_extend2 and the CIGAR choice of its ksw_global2 against
selections made from explicitly enumerated alignment paths, not from any DP recurrence.

orc_enum_prefix walks every alignment prefix from the origin under bwa's gap rules (a gap run
opens only from an aligned pair or the origin, so a deletion run never touches an insertion run;
a run costs o + len*e) and keeps, per cell, the best prefix ending there with an aligned pair, a
deletion or an insertion.  From those values:

* ksw_extend2 (crates/bwa-extend/src/sw.rs:203-493).  With w and end_bonus large (no band), zdrop 0
  and h0 large enough that no score reaches 0 (no band tightening, no early exit), the reference's
  H(i, j) is h0 plus the best prefix score at (i, j).  The result is then fixed by the documented
  rules alone: a row's maximum is taken at its LAST column among ties (sw.rs:390,
  `mj = if row_max > h { mj } else { j }`); the best cell moves only to a row whose maximum is
  STRICTLY greater (sw.rs:440); the to-end score takes the LAST row among ties (sw.rs:424,
  `gscore <= h1`).  Each rule, flipped, must disagree with the oracle on some of the jobs, so that a
  flipped rule fails here.
* ksw_global2 (sw.rs:561-740).  The traceback starts in H at the last cell and, walking back,
  takes the aligned pair when M >= E and M >= F, else the deletion when E >= F, else the insertion
  (ties M over E over F, sw.rs:613-616); inside a gap it continues the gap only when continuing is
  STRICTLY better than having opened it from the aligned pair before (sw.rs:624-636).  Walking back
  from the end, ties therefore push gaps to the left.  Hand-worked homopolymer and tandem-repeat
  cases give the expected gap position.
"""
import numpy as np
import pytest

import oracle_binding as ob

NONE = ob.ENUM_NONE
M, I, D = 0, 1, 2


def _hmax(Mx, Ex, Fx):
    H = np.maximum(np.maximum(Mx, Ex), Fx)
    assert (H > NONE).all()
    return H


def extend_by_rules(q, t, mat, o_del, e_del, o_ins, e_ins, h0, col_tie_later=True, row_strict=True,
                    end_tie_later=True):
    """ksw_extend2's six fields selected from enumerated prefix scores by the documented rules (the
    keyword arguments flip one rule each)."""
    Mx, Ex, Fx = ob.enum_prefix(q, t, mat, o_del, e_del, o_ins, e_ins, 1 << 20, h0)
    H = _hmax(Mx, Ex, Fx)
    tlen, qlen = H.shape
    best, bi, bj, max_off = h0, -1, -1, 0
    gscore, gi = -1, -1
    for i in range(tlen):
        row = H[i]
        rmax = int(row.max())
        hits = np.nonzero(row == rmax)[0]
        mj = int(hits[-1] if col_tie_later else hits[0])
        h_end = int(row[qlen - 1])
        if (gscore <= h_end) if end_tie_later else (gscore < h_end):
            gscore, gi = h_end, i
        if (rmax > best) if row_strict else (rmax >= best):
            best, bi, bj = rmax, i, mj
            max_off = max(max_off, abs(mj - i))
    return (best, bj + 1, bi + 1, gi + 1, gscore, max_off)


def _tie_jobs(rng, n, alphabet=2, lmax=7):
    """Small jobs over a small alphabet with planted repeats: many equal-scoring cells."""
    jobs = []
    for _ in range(n):
        unit = rng.integers(0, alphabet, int(rng.integers(1, 4)))
        q = np.resize(unit, int(rng.integers(1, lmax + 1)))
        if rng.random() < 0.5:
            q = rng.integers(0, alphabet, len(q))
        t = list(q)
        for _ in range(int(rng.integers(0, 3))):  # substitutions, insertions, deletions
            k = int(rng.integers(0, len(t) + 1))
            r = rng.random()
            if r < 0.33 and k < len(t):
                t[k] = int(rng.integers(0, alphabet))
            elif r < 0.66:
                t.insert(k, int(rng.integers(0, alphabet)))
            elif len(t) > 1 and k < len(t):
                del t[k]
        t = np.array(t[:lmax] if t else [0], dtype=np.uint8)
        jobs.append((np.asarray(q, dtype=np.uint8), t))
    return jobs


SCORINGS = [(1, 4, 6, 1, 6, 1), (2, 3, 5, 2, 3, 1), (1, 1, 2, 1, 4, 2), (3, 2, 1, 3, 2, 1)]
# a gap costs about a match here, so equal maxima at two columns of a new best row are common
EXT_SCORINGS = SCORINGS + [(2, 1, 1, 1, 1, 1)]


def test_extend2_tie_rules_by_enumeration():
    flips = {"col": 0, "row": 0, "end": 0}
    for a, b, o_del, e_del, o_ins, e_ins in EXT_SCORINGS:
        rng = np.random.default_rng(1000 * a + 100 * b + 10 * o_del + e_ins)
        mat = ob.scmat(a, b)
        for q, t in _tie_jobs(rng, 1500):
            h0 = int(rng.integers(80, 200))  # no prefix of these lengths falls to 0
            want = extend_by_rules(q, t, mat, o_del, e_del, o_ins, e_ins, h0)
            got = ob.extend2(q, t, mat, o_del, e_del, o_ins, e_ins, 1000, 1000, 0, h0)
            assert got == want, ((a, b, o_del, e_del, o_ins, e_ins), list(q), list(t), h0, got, want)
            flips["col"] += extend_by_rules(q, t, mat, o_del, e_del, o_ins, e_ins, h0, col_tie_later=False) != got
            flips["row"] += extend_by_rules(q, t, mat, o_del, e_del, o_ins, e_ins, h0, row_strict=False) != got
            flips["end"] += extend_by_rules(q, t, mat, o_del, e_del, o_ins, e_ins, h0, end_tie_later=False) != got
    # each rule decides some of these jobs: flipping any one of them fails the comparison above
    assert min(flips.values()) >= 10, flips


def global_by_rules(q, t, mat, o_del, e_del, o_ins, e_ins, w, m_first=True, open_on_tie=True):
    """ksw_global2's score and CIGAR from enumerated prefix scores and the reference's traceback
    priority (see the module docstring); m_first=False takes the gaps first on ties, and
    open_on_tie=False continues a gap on ties (each flips one rule)."""
    Mx, Ex, Fx = ob.enum_prefix(q, t, mat, o_del, e_del, o_ins, e_ins, w, 0)
    tlen, qlen = Mx.shape
    H = np.maximum(np.maximum(Mx, Ex), Fx)
    score = int(H[tlen - 1, qlen - 1])
    runs = []

    def push(op, n=1):
        if runs and runs[-1][0] == op:
            runs[-1][1] += n
        else:
            runs.append([op, n])

    i, j, st = tlen - 1, qlen - 1, "H"
    while i >= 0 and j >= 0:
        if st == "H":
            m, e, f = Mx[i, j], Ex[i, j], Fx[i, j]
            if m_first:
                st = "M" if (m >= e and m >= f) else ("E" if e >= f else "F")
            else:
                st = "F" if (f >= e and f >= m) else ("E" if e >= m else "M")
        if st == "M":
            push(M)
            i, j, st = i - 1, j - 1, "H"
        elif st == "E":
            push(D)
            i -= 1
            if i >= 0:
                ce, oe = Ex[i, j] - e_del, Mx[i, j] - (o_del + e_del)
                cont = Ex[i, j] > NONE and (Mx[i, j] == NONE or ce > oe or (ce == oe and not open_on_tie))
                st = "E" if cont else "M"
        else:
            push(I)
            j -= 1
            if j >= 0:
                cf, of = Fx[i, j] - e_ins, Mx[i, j] - (o_ins + e_ins)
                cont = Fx[i, j] > NONE and (Mx[i, j] == NONE or cf > of or (cf == of and not open_on_tie))
                st = "F" if cont else "M"
    if i >= 0:
        push(D, i + 1)
    if j >= 0:
        push(I, j + 1)
    return score, [(n << 4) | op for op, n in reversed(runs)]


# with a free gap open, continuing a gap and opening it anew can tie
GLB_SCORINGS = SCORINGS + [(1, 2, 0, 1, 0, 1)]


def test_global2_cigar_choice_by_enumeration():
    n_gapped = flip_m = flip_open = 0
    for a, b, o_del, e_del, o_ins, e_ins in GLB_SCORINGS:
        rng = np.random.default_rng(7 + 1000 * a + 100 * b + 10 * o_del + e_ins)
        mat = ob.scmat(a, b)
        for q, t in _tie_jobs(rng, 800, alphabet=3 if o_del == 0 else 2, lmax=6):
            d = abs(len(q) - len(t))
            for w in (d, d + 1, 6):
                want = global_by_rules(q, t, mat, o_del, e_del, o_ins, e_ins, w)
                got = ob.global2(q, t, mat, o_del, e_del, o_ins, e_ins, w)
                assert got == want, ((a, b, o_del, e_del, o_ins, e_ins), list(q), list(t), w, got, want)
                n_gapped += any((c & 0xf) != M for c in got[1])
                flip_m += global_by_rules(q, t, mat, o_del, e_del, o_ins, e_ins, w, m_first=False) != got
                flip_open += global_by_rules(q, t, mat, o_del, e_del, o_ins, e_ins, w, open_on_tie=False) != got
    assert n_gapped > 1000
    # both rules decide CIGARs here: either one flipped fails the comparison above
    assert flip_m >= 100 and flip_open >= 5, (flip_m, flip_open)


def _cig(s):
    """'1M2I3M' -> CIGAR words."""
    out, num = [], ""
    for ch in s:
        if ch.isdigit():
            num += ch
        else:
            out.append((int(num) << 4) | {"M": M, "I": I, "D": D}[ch])
            num = ""
    return out


A_, C_, G_, T_ = 0, 1, 2, 3


@pytest.mark.parametrize("q,t,w,want", [
    # a homopolymer deletion: every position scores the same; the traceback's M-first ties push it
    # to the left end, a leading deletion
    ([A_] * 4, [A_] * 5, 2, "1D4M"),
    ([A_] * 5, [A_] * 4, 2, "1I4M"),
    # the same inside flanks: leftmost base of the run
    ([G_, A_, A_, A_, A_, T_], [G_, A_, A_, A_, T_], 2, "1M1I4M"),
    ([G_, A_, A_, A_, T_], [G_, A_, A_, A_, A_, T_], 2, "1M1D4M"),
    # a tandem-repeat unit: leftmost copy
    ([T_, A_, C_, A_, C_, G_], [T_, A_, C_, G_], 3, "1M2I3M"),
    ([T_, A_, C_, G_], [T_, A_, C_, A_, C_, G_], 3, "1M2D3M"),
    ([A_, C_, A_, C_, A_, C_], [A_, C_, A_, C_], 2, "2I4M"),
])
def test_global2_hand_worked_gap_positions(q, t, w, want):
    mat = ob.scmat(1, 4)
    score, cig = ob.global2(q, t, mat, 6, 1, 6, 1, w)
    assert cig == _cig(want), (cig, _cig(want))
    assert (score, cig) == global_by_rules(np.array(q, np.uint8), np.array(t, np.uint8), mat, 6, 1, 6, 1, w)
    n_gap = sum(c >> 4 for c in cig if (c & 0xf) != M)
    assert score == min(len(q), len(t)) - (6 + n_gap)
