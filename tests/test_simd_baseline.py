"""The SIMD CPU baseline (oracle/ksw_simd.c) equals the scalar oracle on every output field.

The baseline is the reference's inter-sequence form (16 jobs in the lanes of an AVX2 register, the
union band with per-lane masks, a scalar per-lane row epilogue; batched.rs:1370-1841,
matesw.rs:752-877) that bench.py times beside the GPU.  These tests run it on the reference's
fixed-seed gate sweeps (single jobs over swept scorings, ragged batches for lane cross-talk,
read-sized and ungapped-mode batches, the rescue gate under both scorings) and on the synthetic
workloads, against the scalar ksw_extend2 / batched ksw_align2.
"""
import numpy as np
import pytest

import bwg
import lcg_gates as G
import oracle_binding as ob


def _batch(queries, targets, h0s=None):
    return bwg.JobBatch.from_pairs(queries, targets, h0s)


def test_simd_extend_single_gate():
    for c in G.ext_single_cases():
        b = _batch([c["q"]], [c["t"]], [c["h0"]])
        args = (c["mat"], c["o_del"], c["e_del"], c["o_ins"], c["e_ins"], c["w"], c["end_bonus"], c["zdrop"])
        got = ob.simd_extend_batch(b, *args, nthreads=1)
        want = ob.extend_batch(b, *args, nthreads=1)
        assert np.array_equal(got, want), (c, got, want)


@pytest.mark.parametrize("gen", [G.ext_batch_rounds, G.ext_read_sized, G.ext_ungapped])
def test_simd_extend_batch_gates(gen):
    for c in gen():
        b = _batch(c["queries"], c["targets"], c["h0s"])
        args = (c["mat"], c["o_del"], c["e_del"], c["o_ins"], c["e_ins"], c["w"], c["end_bonus"], c["zdrop"])
        got = ob.simd_extend_batch(b, *args, nthreads=2)
        assert np.array_equal(got, ob.extend_batch(b, *args, nthreads=1))


@pytest.mark.parametrize("kind", [bwg.WORKLOAD_EXT_TAIL, bwg.WORKLOAD_EXT_SMALL])
def test_simd_extend_workloads(kind):
    b = bwg.generate(kind, 30000, seed=11)
    mat = ob.scmat(1, 4)
    got = ob.simd_extend_batch(b, mat, 6, 1, 6, 1, 100, 5, 100, nthreads=4)
    assert np.array_equal(got, ob.extend_batch(b, mat, 6, 1, 6, 1, 100, 5, 100, nthreads=4))


def test_simd_extend_outside_its_range_runs_scalar():
    # a non-uniform matrix and scores past the 16-bit lanes take the scalar path, same results
    b = bwg.generate(bwg.WORKLOAD_EXT_TAIL, 3000, seed=12)
    mat = ob.scmat(1, 4).copy()
    mat[6] = 2  # C against C scores 2: not bwa's uniform matrix
    got = ob.simd_extend_batch(b, mat, 6, 1, 6, 1, 100, 5, 100, nthreads=2)
    assert np.array_equal(got, ob.extend_batch(b, mat, 6, 1, 6, 1, 100, 5, 100, nthreads=2))
    big = _batch([[0, 1, 2, 3] * 40] * 20, [[0, 1, 2, 3] * 50] * 20, [29900] * 20)
    mat = ob.scmat(1, 4)
    got = ob.simd_extend_batch(big, mat, 6, 1, 6, 1, 100, 5, 100, nthreads=1)
    assert np.array_equal(got, ob.extend_batch(big, mat, 6, 1, 6, 1, 100, 5, 100, nthreads=1))


@pytest.mark.parametrize("a,b,minsc", G.RESCUE_SCORINGS)
def test_simd_rescue_gate(a, b, minsc):
    qs, ts = G.rescue_jobs()
    bb = _batch(qs, ts)
    mat = ob.scmat(a, b)
    got = ob.simd_align2_batch(bb, mat, 6, 1, 6, 1, minsc, a, nthreads=3)
    want = ob.align2_batch(bb, mat, 6, 1, 6, 1, minsc, a, batched=True, nthreads=3)
    assert np.array_equal(got, want)
    assert (want["score2"] >= 0).sum() > 50 and (want["qb"] >= 0).sum() > 500


def test_simd_rescue_workload_and_ragged_chunks():
    b = bwg.generate(bwg.WORKLOAD_RESCUE, 4000, seed=13)
    mat = ob.scmat(1, 4)
    want = ob.align2_batch(b, mat, 6, 1, 6, 1, 20, 1, nthreads=4)
    assert np.array_equal(ob.simd_align2_batch(b, mat, 6, 1, 6, 1, 20, 1, nthreads=4), want)
    for n in (1, 15, 17, 33):
        sub = b.subset(np.arange(n))
        assert np.array_equal(ob.simd_align2_batch(sub, mat, 6, 1, 6, 1, 20, 1, nthreads=1), want[:n])
    # empty sides give the default result
    e = _batch([[], [0, 1, 2]], [[1, 2, 3], []])
    got = ob.simd_align2_batch(e, mat, 6, 1, 6, 1, 20, 1, nthreads=1)
    assert np.array_equal(got, ob.align2_batch(e, mat, 6, 1, 6, 1, 20, 1, nthreads=1))
