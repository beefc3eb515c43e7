"""GPU tests of the tie rules: the extension and global-alignment kernels on the tie-heavy jobs of
test_oracle_ties.py, compared with the selections made there from explicitly enumerated alignment
paths (not only with the oracle)."""
import sys

import numpy as np
import pytest

import bwg
import oracle_binding as ob
from test_oracle_ties import EXT_SCORINGS, GLB_SCORINGS, _tie_jobs, extend_by_rules, global_by_rules

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


@pytest.mark.parametrize("form", [bwg.FORM_BATCHED, bwg.FORM_PLAIN], ids=["batched", "plain"])
def test_extension_ties_on_device(gpu_ctx, form):
    _say("extension ties form=%d" % form)
    gpu_ctx.set_form(form)
    for a, b, o_del, e_del, o_ins, e_ins in EXT_SCORINGS:
        rng = np.random.default_rng(500 + 1000 * a + 100 * b + 10 * o_del + e_ins)
        jobs = _tie_jobs(rng, 1500)
        h0 = rng.integers(80, 200, len(jobs))
        batch = bwg.JobBatch.from_pairs([q for q, _ in jobs], [t for _, t in jobs], h0)
        mat = ob.scmat(a, b)
        got = gpu_ctx.extend_arrays(batch, bwg.ext_params(5, mat, o_del, e_del, o_ins, e_ins, 1000, 1000, 0))
        for k, (q, t) in enumerate(jobs):
            want = extend_by_rules(q, t, mat, o_del, e_del, o_ins, e_ins, int(h0[k]))
            assert tuple(int(x) for x in got[k]) == want, ((a, b, o_del, e_del, o_ins, e_ins), k)
    gpu_ctx.set_form(bwg.FORM_BATCHED)


def test_global_cigar_ties_on_device(gpu_ctx):
    _say("global ties")
    gpu_ctx.set_form(bwg.FORM_BATCHED)
    for a, b, o_del, e_del, o_ins, e_ins in GLB_SCORINGS:
        rng = np.random.default_rng(900 + 1000 * a + 100 * b + 10 * o_del + e_ins)
        jobs = _tie_jobs(rng, 600, alphabet=3 if o_del == 0 else 2, lmax=6)
        qs, ts, ws = [], [], []
        for q, t in jobs:
            d = abs(len(q) - len(t))
            for w in (d, d + 1, 6):
                qs.append(q)
                ts.append(t)
                ws.append(w)
        batch = bwg.JobBatch.from_pairs(qs, ts)
        mat = ob.scmat(a, b)
        score, ncig, cig = gpu_ctx.global_arrays(batch, bwg.global_params(5, mat, o_del, e_del, o_ins, e_ins),
                                                 np.array(ws))
        off = 0
        for k in range(len(qs)):
            want = global_by_rules(qs[k], ts[k], mat, o_del, e_del, o_ins, e_ins, ws[k])
            got = (int(score[k]), [int(x) for x in cig[off:off + int(ncig[k])]])
            off += int(ncig[k])
            assert got == want, ((a, b, o_del, e_del, o_ins, e_ins), list(qs[k]), list(ts[k]), ws[k], got, want)
