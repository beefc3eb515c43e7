"""GPU parity under random scorings: batched extension and rescue, bit-exact to the CPU oracle.

The gate sweeps and workloads fix a few scorings; here each case draws its own match and mismatch
scores, gap penalties, band width, z-drop and end bonus, and a mixed batch of short, medium and
long queries (exact, lightly and heavily mutated, with indels, random targets, N bases, targets of
0-3 bases).  Every batch runs through one staged batch twice, so the second run re-reads
whatever the first left on the device (queues, compacted bins, learned shares), and both runs must
equal the oracle on every output field.  The rescue cases plant a mutated copy of each query in a
longer window (or none), under random scores, gap penalties and minimum scores, and call the
batched ksw_align2 twice.
"""
import sys

import numpy as np
import pytest

import bwg
import oracle_binding as ob

pytestmark = pytest.mark.gpu


def _mixed_batch(rng, n):
    qs, ts, h0 = [], [], []
    for _ in range(n):
        ql = int(rng.choice([rng.integers(1, 33), rng.integers(33, 140), rng.integers(140, 300)]))
        q = rng.integers(0, 5 if rng.random() < 0.05 else 4, ql).astype(np.uint8)
        t = q.copy()
        m = rng.random(ql) < rng.choice([0.0, 0.02, 0.1, 0.4])
        t[m] = (t[m] + 1) % 4
        if ql > 6 and rng.random() < 0.3:
            p = int(rng.integers(1, ql - 1))
            if rng.random() < 0.5:
                t = np.delete(t, p)
            else:
                t = np.insert(t, p, rng.integers(0, 4, int(rng.integers(1, 6))).astype(np.uint8))
        if rng.random() < 0.2:
            t = rng.integers(0, 4, max(1, len(t))).astype(np.uint8)
        t = np.concatenate([t, rng.integers(0, 4, int(rng.integers(0, 120))).astype(np.uint8)])
        if rng.random() < 0.02:
            t = t[:int(rng.integers(0, 4))]
        qs.append(q)
        ts.append(t)
        h0.append(int(rng.integers(1, 260)))
    return bwg.JobBatch.from_pairs(qs, ts, np.asarray(h0, dtype=np.int32))


@pytest.mark.parametrize("seed", range(10))
def test_batched_extension_random_scoring(gpu_ctx, seed):
    rng = np.random.default_rng(9100 + seed)
    o_del, e_del = int(rng.integers(0, 12)), int(rng.integers(1, 6))
    o_ins, e_ins = int(rng.integers(0, 12)), int(rng.integers(1, 6))
    a, b = int(rng.integers(1, 4)), int(rng.integers(1, 7))
    w = int(rng.choice([1, 3, 7, 20, 50, 100, 150, 250]))
    zdrop = int(rng.choice([0, 10, 50, 100, 200]))
    eb = int(rng.integers(0, 12))
    print("seed %d: a=%d b=%d o_del=%d e_del=%d o_ins=%d e_ins=%d w=%d zdrop=%d end_bonus=%d"
          % (seed, a, b, o_del, e_del, o_ins, e_ins, w, zdrop, eb), file=sys.stderr, flush=True)
    mat = ob.scmat(a, b)
    batch = _mixed_batch(rng, 3000)
    want = ob.extend_batch(batch, mat, o_del, e_del, o_ins, e_ins, w, eb, zdrop, nthreads=8)
    gpu_ctx.set_form(bwg.FORM_BATCHED)
    st = gpu_ctx.stage(batch, bwg.ext_params(5, mat, o_del, e_del, o_ins, e_ins, w, eb, zdrop))
    try:
        for run in range(2):
            st.run()
            got = st.results()
            bad = np.flatnonzero(got != want)
            assert bad.size == 0, "run %d: %d jobs differ, first %d: got %s want %s" % (
                run, bad.size, bad[0], got[bad[0]], want[bad[0]])
    finally:
        st.close()


def _rescue_batch(rng, n):
    qs, ts = [], []
    for _ in range(n):
        ql = int(rng.integers(20, 201))
        q = rng.integers(0, 4, ql).astype(np.uint8)
        tl = int(rng.integers(ql // 2, 601))
        t = rng.integers(0, 5 if rng.random() < 0.05 else 4, tl).astype(np.uint8)
        if rng.random() < 0.8:
            c = q.copy()
            m = rng.random(ql) < rng.choice([0.0, 0.03, 0.1, 0.3])
            c[m] = (c[m] + 1) % 4
            if ql > 10 and rng.random() < 0.3:
                p = int(rng.integers(1, ql - 1))
                c = np.delete(c, p) if rng.random() < 0.5 else np.insert(c, p, rng.integers(0, 4, 3).astype(np.uint8))
            at = int(rng.integers(0, max(1, tl - len(c) + 1)))
            t = np.concatenate([t[:at], c, t[at + len(c):]]).astype(np.uint8)
        qs.append(q)
        ts.append(t)
    return bwg.JobBatch.from_pairs(qs, ts)


@pytest.mark.parametrize("seed", range(6))
def test_batched_rescue_random_scoring(gpu_ctx, seed):
    rng = np.random.default_rng(9300 + seed)
    o_del, e_del = int(rng.integers(0, 12)), int(rng.integers(1, 6))
    o_ins, e_ins = int(rng.integers(0, 12)), int(rng.integers(1, 6))
    a, b = int(rng.integers(1, 4)), int(rng.integers(1, 7))
    minsc = int(a * rng.integers(5, 31))
    print("seed %d: a=%d b=%d o_del=%d e_del=%d o_ins=%d e_ins=%d minsc=%d"
          % (seed, a, b, o_del, e_del, o_ins, e_ins, minsc), file=sys.stderr, flush=True)
    mat = ob.scmat(a, b)
    batch = _rescue_batch(rng, 1500)
    want = ob.align2_batch(batch, mat, o_del, e_del, o_ins, e_ins, minsc, a)
    gpu_ctx.set_form(bwg.FORM_BATCHED)
    for run in range(2):
        got = gpu_ctx.align2_arrays(batch, bwg.ksw_params(5, mat, o_del, e_del, o_ins, e_ins, minsc, a))
        bad = np.flatnonzero(got != want)
        assert bad.size == 0, "run %d: %d jobs differ, first %d: got %s want %s" % (
            run, bad.size, bad[0], got[bad[0]], want[bad[0]])
