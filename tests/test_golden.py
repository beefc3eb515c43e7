"""CPU: the oracle and the workload generators reproduce the committed golden fixtures."""
import numpy as np
import pytest

import bwg
import golden_fixtures as gf
import lcg_gates as G
import oracle_binding as ob


@pytest.mark.parametrize("prefix,kind", [("small_", bwg.WORKLOAD_EXT_SMALL), ("tail_", bwg.WORKLOAD_EXT_TAIL)])
def test_extension_workload_fixture(prefix, kind):
    d = gf.load("ext_workloads.npz")
    b = gf.batch(d, prefix)
    regen = bwg.generate(kind, 400, seed=42)  # the generator is deterministic across builds
    for name in ("arena", "q_off", "q_len", "t_off", "t_len", "h0"):
        assert np.array_equal(getattr(regen, name), getattr(b, name)), name
    got = gf.as2d(ob.extend_batch(b, bwg.default_mat(), 6, 1, 6, 1, 100, 5, 100), gf.EXT_FIELDS)
    assert np.array_equal(got, d[prefix + "want"])


def test_extension_scoring_sweep_fixture():
    d = gf.load("ext_scorings.npz")
    b = gf.batch(d)
    mat = G.scmat(1, 4)
    for sc, idx in gf.scoring_groups(d):
        got = gf.as2d(ob.extend_batch(b.subset(idx), mat, *sc), gf.EXT_FIELDS)
        assert np.array_equal(got, d["want"][idx]), sc


def test_rescue_fixtures():
    d = gf.load("rescue.npz")
    b = gf.batch(d, "work_")
    regen = bwg.generate(bwg.WORKLOAD_RESCUE, 100, seed=42)
    assert np.array_equal(regen.arena, b.arena) and np.array_equal(regen.t_len, b.t_len)
    got = gf.as2d(ob.align2_batch(b, bwg.default_mat(), 6, 1, 6, 1, 20, 1), gf.KSW_FIELDS)
    assert np.array_equal(got, d["work_want"])
    g = gf.batch(d, "gate_")
    for a, bb, minsc in G.RESCUE_SCORINGS:
        want = d["gate_want_%d_%d_%d" % (a, bb, minsc)]
        for batched in (True, False):  # batched_ksw_align2 and scalar ksw_align2 semantics agree
            got = gf.as2d(ob.align2_batch(g, G.scmat(a, bb), 6, 1, 6, 1, minsc, a, batched=batched),
                          gf.KSW_FIELDS)
            assert np.array_equal(got, want), (a, batched)
