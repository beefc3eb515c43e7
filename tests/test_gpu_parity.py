"""GPU parity: every output field bit-exact to the CPU oracle, through the C ABI.

Both execution forms are run on the reference's own acceptance-gate sweeps, on samples of the
synthetic workloads, on edge cases, and on size-independent properties at full batch sizes.
"""
import os
import sys

import numpy as np
import pytest

import bwg
import lcg_gates as G
import oracle_binding as ob

pytestmark = pytest.mark.gpu
FORMS = [bwg.FORM_PLAIN, bwg.FORM_BATCHED]
FORM_IDS = ["plain", "batched"]


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


def _ext(ctx, form, batch, mat, o_del, e_del, o_ins, e_ins, w, eb, zd):
    ctx.set_form(form)
    before = ctx.stat(bwg.STAT_LAUNCHES)
    out = ctx.extend_arrays(batch, bwg.ext_params(5, mat, o_del, e_del, o_ins, e_ins, w, eb, zd))
    assert len(batch) == 0 or ctx.stat(bwg.STAT_LAUNCHES) > before  # the device really ran
    return out


def _ksw(ctx, form, batch, mat, o_del, e_del, o_ins, e_ins, minsc, max_sc):
    ctx.set_form(form)
    before = ctx.stat(bwg.STAT_LAUNCHES)
    out = ctx.align2_arrays(batch, bwg.ksw_params(5, mat, o_del, e_del, o_ins, e_ins, minsc, max_sc))
    assert len(batch) == 0 or ctx.stat(bwg.STAT_LAUNCHES) > before
    return out


def _tuples(arr):
    return [tuple(int(x) for x in r) for r in arr]


@pytest.mark.parametrize("form", FORMS, ids=FORM_IDS)
def test_ext_single_gate(gpu_ctx, form):
    """assert_backend_matches_scalar (crates/bwa-extend/src/lib.rs:283-354): 2000 cases, each with
    its own scoring, run as one-job batches grouped by scoring."""
    _say("ext_single_gate form=%d" % form)
    cases = list(G.ext_single_cases())
    for k, c in enumerate(cases):
        b = bwg.JobBatch.from_pairs([c["q"]], [c["t"]], [c["h0"]])
        got = _tuples(_ext(gpu_ctx, form, b, c["mat"], c["o_del"], c["e_del"], c["o_ins"], c["e_ins"],
                           c["w"], c["end_bonus"], c["zdrop"]))[0]
        want = ob.extend2(c["q"], c["t"], c["mat"], c["o_del"], c["e_del"], c["o_ins"], c["e_ins"],
                          c["w"], c["end_bonus"], c["zdrop"], c["h0"])
        assert got == want, "case %d" % k


@pytest.mark.parametrize("form", FORMS, ids=FORM_IDS)
@pytest.mark.parametrize("gen", [G.ext_batch_rounds, G.ext_read_sized, G.ext_ungapped],
                         ids=["batch_rounds", "read_sized", "ungapped"])
def test_ext_batch_gates(gpu_ctx, form, gen):
    _say("ext_batch_gate %s form=%d" % (gen.__name__, form))
    for r, c in enumerate(gen()):
        b = bwg.JobBatch.from_pairs(c["queries"], c["targets"], c["h0s"])
        args = (c["mat"], c["o_del"], c["e_del"], c["o_ins"], c["e_ins"], c["w"], c["end_bonus"], c["zdrop"])
        got = _tuples(_ext(gpu_ctx, form, b, *args))
        for k in range(len(b)):
            assert got[k] == ob.extend2(c["queries"][k], c["targets"][k], *args, c["h0s"][k]), \
                "round %d job %d" % (r, k)


@pytest.mark.parametrize("form", FORMS, ids=FORM_IDS)
@pytest.mark.parametrize("a,b,minsc", G.RESCUE_SCORINGS)
def test_rescue_gate(gpu_ctx, form, a, b, minsc):
    """matesw_equals_scalar (crates/bwa-neon/src/matesw.rs:5045-5139), both scorings."""
    _say("rescue_gate form=%d a=%d" % (form, a))
    qs, ts = G.rescue_jobs()
    batch = bwg.JobBatch.from_pairs(qs, ts)
    mat = ob.scmat(a, b)
    got = _ksw(gpu_ctx, form, batch, mat, 6, 1, 6, 1, minsc, a)
    want = ob.align2_batch(batch, mat, 6, 1, 6, 1, minsc, a)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, "first mismatch job %d: got %s want %s" % (bad[0], got[bad[0]], want[bad[0]])


@pytest.mark.parametrize("form", FORMS, ids=FORM_IDS)
@pytest.mark.parametrize("kind,n", [(bwg.WORKLOAD_EXT_SMALL, 20000), (bwg.WORKLOAD_EXT_TAIL, 30000)],
                         ids=["ext_small", "ext_tail"])
def test_ext_workload_sample(gpu_ctx, form, kind, n):
    _say("ext_workload kind=%d n=%d form=%d" % (kind, n, form))
    batch = bwg.generate(kind, n, seed=42)
    mat = bwg.default_mat()
    got = _ext(gpu_ctx, form, batch, mat, 6, 1, 6, 1, 100, 5, 100)
    want = ob.extend_batch(batch, mat, 6, 1, 6, 1, 100, 5, 100)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, "first mismatch job %d: got %s want %s" % (bad[0], got[bad[0]], want[bad[0]])


@pytest.mark.parametrize("form", FORMS, ids=FORM_IDS)
def test_rescue_workload_sample(gpu_ctx, form):
    _say("rescue_workload form=%d" % form)
    batch = bwg.generate(bwg.WORKLOAD_RESCUE, 4000, seed=42)
    mat = bwg.default_mat()
    # the rescue workload's own minsc (min_seed_len 20 * a)
    got = _ksw(gpu_ctx, form, batch, mat, 6, 1, 6, 1, 20, 1)
    want = ob.align2_batch(batch, mat, 6, 1, 6, 1, 20, 1)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, "first mismatch job %d: got %s want %s" % (bad[0], got[bad[0]], want[bad[0]])
    assert np.mean(got["qb"] >= 0) > 0.5  # the reverse pass ran for the placed mates


@pytest.mark.parametrize("form", FORMS, ids=FORM_IDS)
def test_edge_cases(gpu_ctx, form):
    mat = bwg.default_mat()
    qs = [[0], [4], [0, 1, 2, 3], [1] * 250, [2] * 7, [4] * 9, [0, 1] * 60, [3]]
    ts = [[], [4], [0, 1, 2, 3] + [1] * 200, [1] * 350, [], [4] * 9, [1, 0] * 70, [3] * 3000]
    h0 = [1, 5, 1, 250, 40, 3, 30, 1000]
    b = bwg.JobBatch.from_pairs(qs, ts, h0)
    got = _tuples(_ext(gpu_ctx, form, b, mat, 6, 1, 6, 1, 100, 5, 100))
    for k in range(len(qs)):
        assert got[k] == ob.extend2(qs[k], ts[k], mat, 6, 1, 6, 1, 100, 5, 100, h0[k]), k
    # rescue: empty query, empty target, all-N, exact copy, query longer than target
    qs = [[], [0, 1, 2], [4] * 20, [0, 1, 2, 3] * 40, [2, 3] * 100]
    ts = [[0, 1, 2], [], [4] * 50, [1] * 30 + [0, 1, 2, 3] * 40 + [2] * 30, [2, 3] * 20]
    b = bwg.JobBatch.from_pairs(qs, ts)
    got = _tuples(_ksw(gpu_ctx, form, b, mat, 6, 1, 6, 1, 19, 1))
    for k in range(len(qs)):
        assert got[k] == ob.batched_align2_job(qs[k], ts[k], mat, 6, 1, 6, 1, 19, 1), k


def test_batch_composition_independence(gpu_ctx):
    """A job's result does not depend on what else is in the batch (lane cross-talk)."""
    batch = bwg.generate(bwg.WORKLOAD_EXT_TAIL, 4000, seed=9)
    mat = bwg.default_mat()
    whole = _ext(gpu_ctx, bwg.FORM_BATCHED, batch, mat, 6, 1, 6, 1, 100, 5, 100)
    rng = np.random.default_rng(0)
    idx = rng.permutation(len(batch))[:700]
    part = _ext(gpu_ctx, bwg.FORM_BATCHED, batch.subset(idx), mat, 6, 1, 6, 1, 100, 5, 100)
    assert np.array_equal(part, whole[idx])
    for k in idx[:5]:
        one = _ext(gpu_ctx, bwg.FORM_BATCHED, batch.subset([k]), mat, 6, 1, 6, 1, 100, 5, 100)
        assert np.array_equal(one[0], whole[k])


def test_invalid_jobs_rejected(gpu_ctx):
    mat = bwg.default_mat()
    p = bwg.ext_params(5, mat, 6, 1, 6, 1, 100, 5, 100)
    cases = [
        (bwg.JobBatch.from_pairs([[0, 1]], [[0, 1]], [0]), p),                     # h0 <= 0
        (bwg.JobBatch.from_pairs([[]], [[0, 1]], [5]), p),                         # qlen == 0
        (bwg.JobBatch.from_pairs([[0, 7]], [[0, 1]], [5]), p),                     # code > 4
        (bwg.JobBatch.from_pairs([[0, 1]], [[0, 1]], [5]), bwg.ext_params(5, mat, 6, 0, 6, 1, 100, 5, 100)),
        (bwg.JobBatch.from_pairs([[0, 1]], [[0, 1]], [5]), bwg.ext_params(4, mat, 6, 1, 6, 1, 100, 5, 100)),
    ]
    for b, params in cases:
        with pytest.raises(bwg.BwgError) as e:
            gpu_ctx.extend_arrays(b, params)
        assert e.value.code == bwg.BWG_EINVAL
    bad = bwg.JobBatch.from_pairs([[0, 1]], [[0, 1]], [5])
    bad.q_off[0] = 10**9  # outside the arena
    with pytest.raises(bwg.BwgError):
        gpu_ctx.extend_arrays(bad, p)
    with pytest.raises(bwg.BwgError) as e:
        gpu_ctx.align2_arrays(bwg.JobBatch.from_pairs([[0]], [[0]]), bwg.ksw_params(5, mat, 6, 1, 6, 1, 19, 0))
    assert e.value.code == bwg.BWG_EINVAL


def test_backend_mirror_api(gpu_ctx):
    be = bwg.HipBackend(0)
    assert be.name() == "hip-gfx950"
    mat = bwg.default_mat()
    s = [0, 1, 2, 3] * 4
    assert tuple(be.extend(s, s, 5, mat, 6, 1, 6, 1, 100, 0, 100, 1))[:3] == (17, 16, 16)
    jobs = [bwg.ExtendJob(s, s, 1), bwg.ExtendJob([0] * 10 + [1] * 40, [0] * 10 + [2] * 40, 1)]
    out = be.extend_batch(jobs, 5, mat, 6, 1, 6, 1, 100, 0, 100)
    assert out[0].score == 17 and out[1].score == 11 and out[1].tle == 10
    r = be.batched_ksw_align2([bwg.KswJob([0, 1, 2, 3, 0, 1, 2, 3], [0, 1, 2, 3, 0, 1, 2, 3])],
                              5, mat, 6, 1, 6, 1, 4, 1)
    assert (r[0].score, r[0].qb, r[0].qe, r[0].tb, r[0].te) == (8, 0, 7, 0, 7)
    assert be.extend_batch([], 5, mat, 6, 1, 6, 1, 100, 0, 100) == []


@pytest.mark.parametrize("kind", ["ext", "resc"])
def test_graph_replays_match_eager(gpu_ctx, kind):
    """The first batched run launches eagerly; later runs replay a captured hipGraph.  Every replay,
    with other eager device work and synchronisations in between, must equal the eager result and
    the oracle."""
    _say("graph_replays %s" % kind)
    mat = bwg.default_mat()
    if kind == "ext":
        batch = bwg.generate(bwg.WORKLOAD_EXT_TAIL, 20000, seed=5)
        params = bwg.ext_params(5, mat, 6, 1, 6, 1, 100, 5, 100)
        want = ob.extend_batch(batch, mat, 6, 1, 6, 1, 100, 5, 100)
        other = bwg.generate(bwg.WORKLOAD_EXT_SMALL, 3000, seed=6)
        oparams = params
    else:
        batch = bwg.generate(bwg.WORKLOAD_RESCUE, 3000, seed=5)
        params = bwg.ksw_params(5, mat, 6, 1, 6, 1, 19, 1)
        want = ob.align2_batch(batch, mat, 6, 1, 6, 1, 19, 1)
        other = bwg.generate(bwg.WORKLOAD_EXT_SMALL, 3000, seed=6)
        oparams = bwg.ext_params(5, mat, 6, 1, 6, 1, 100, 5, 100)
    gpu_ctx.set_form(bwg.FORM_BATCHED)
    st = gpu_ctx.stage(batch, params)
    results = []
    for r in range(5):
        st.run()
        results.append(st.results())
        gpu_ctx.extend_arrays(other, oparams)  # unrelated eager work between replays
    st.close()
    for r, got in enumerate(results):
        bad = np.nonzero(got != want)[0]
        assert bad.size == 0, "run %d: first mismatch job %d" % (r, bad[0])


def test_serial_and_concurrent_bins_agree(gpu_ctx):
    mat = bwg.default_mat()
    batch = bwg.generate(bwg.WORKLOAD_EXT_TAIL, 20000, seed=8)
    params = bwg.ext_params(5, mat, 6, 1, 6, 1, 100, 5, 100)
    gpu_ctx.set_form(bwg.FORM_BATCHED)
    outs = []
    for conc in (0, 1):
        gpu_ctx.set_option(bwg.OPT_CONCURRENT_BINS, conc)
        outs.append(gpu_ctx.extend_arrays(batch, params))
    gpu_ctx.set_option(bwg.OPT_CONCURRENT_BINS, 1)
    assert np.array_equal(outs[0], outs[1])
    assert np.array_equal(outs[0], ob.extend_batch(batch, mat, 6, 1, 6, 1, 100, 5, 100))


def test_ungapped_fast_path_on_and_off(gpu_ctx):
    """The ungapped-diagonal pre-pass (crates/bwa-neon/src/batched.rs:602-697) resolves most clean
    extensions without the DP; results are the oracle's either way, and its hit count equals the
    oracle restatement's."""
    _say("ungapped fast path on/off")
    mat = bwg.default_mat()
    batch = bwg.generate(bwg.WORKLOAD_EXT_SMALL, 30000, seed=17)
    params = bwg.ext_params(5, mat, 6, 1, 6, 1, 100, 5, 100)
    want = ob.extend_batch(batch, mat, 6, 1, 6, 1, 100, 5, 100)
    _, hit = ob.ungapped_batch(batch, mat, 6, 1, 6, 1, 100)
    gpu_ctx.set_form(bwg.FORM_BATCHED)
    try:
        for on in (1, 0):
            gpu_ctx.set_option(bwg.OPT_UNGAPPED, on)
            st = gpu_ctx.stage(batch, params)
            st.run()
            got = st.results()
            hits = st.ungapped_hits()
            st.close()
            bad = np.nonzero(got != want)[0]
            assert bad.size == 0, (on, int(bad[0]))
            assert hits == (int(hit.sum()) if on else 0)
    finally:
        gpu_ctx.set_option(bwg.OPT_UNGAPPED, 1)


@pytest.mark.parametrize("form", FORMS, ids=FORM_IDS)
def test_large_insertion_extension_penalty(gpu_ctx, form):
    # e_ins = 5e6: the group kernels' F scan would carry 16 blocks of 33 columns of it past int32,
    # so the batched form runs these jobs in the plain kernel; the gap bound (e_ins * (len + 1)
    # below 2^30) keeps the jobs under 200 columns.  Every field equals the oracle's.
    _say("large insertion extension penalty, %s" % FORM_IDS[FORMS.index(form)])
    rng = np.random.default_rng(97)
    qs, ts = [], []
    for _ in range(3000):
        q = rng.integers(0, 4, int(rng.integers(1, 150))).astype(np.uint8)
        t = np.concatenate([q, rng.integers(0, 4, int(rng.integers(0, 40))).astype(np.uint8)])
        if len(q) > 4:
            t = np.delete(t, int(rng.integers(1, len(q) - 1)))           # a deletion
            t = np.insert(t, int(rng.integers(1, len(q) - 1)), 2)       # an insertion
        qs.append(q)
        ts.append(t)
    batch = bwg.JobBatch.from_pairs(qs, ts, rng.integers(19, 150, len(qs)).astype(np.int32))
    mat = bwg.default_mat()
    got = _ext(gpu_ctx, form, batch, mat, 6, 1, 6, 5_000_000, 100, 5, 100)
    want = ob.extend_batch(batch, mat, 6, 1, 6, 5_000_000, 100, 5, 100)
    assert np.array_equal(got, want)
