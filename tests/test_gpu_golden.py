"""GPU: both execution forms reproduce the committed golden fixtures through the C ABI."""
import sys

import numpy as np
import pytest

import bwg
import golden_fixtures as gf
import lcg_gates as G

pytestmark = pytest.mark.gpu
FORMS = [bwg.FORM_PLAIN, bwg.FORM_BATCHED]
FORM_IDS = ["plain", "batched"]


@pytest.mark.parametrize("form", FORMS, ids=FORM_IDS)
def test_golden_extension(gpu_ctx, form):
    print("golden extension form=%d" % form, file=sys.stderr, flush=True)
    gpu_ctx.set_form(form)
    d = gf.load("ext_workloads.npz")
    p = bwg.ext_params(5, bwg.default_mat(), 6, 1, 6, 1, 100, 5, 100)
    for prefix in ("small_", "tail_"):
        got = gf.as2d(gpu_ctx.extend_arrays(gf.batch(d, prefix), p), gf.EXT_FIELDS)
        assert np.array_equal(got, d[prefix + "want"]), prefix
    d = gf.load("ext_scorings.npz")
    b = gf.batch(d)
    mat = G.scmat(1, 4)
    for sc, idx in gf.scoring_groups(d):
        got = gf.as2d(gpu_ctx.extend_arrays(b.subset(idx), bwg.ext_params(5, mat, *sc)), gf.EXT_FIELDS)
        assert np.array_equal(got, d["want"][idx]), sc


@pytest.mark.parametrize("form", FORMS, ids=FORM_IDS)
def test_golden_rescue(gpu_ctx, form):
    print("golden rescue form=%d" % form, file=sys.stderr, flush=True)
    gpu_ctx.set_form(form)
    d = gf.load("rescue.npz")
    got = gf.as2d(gpu_ctx.align2_arrays(gf.batch(d, "work_"),
                                        bwg.ksw_params(5, bwg.default_mat(), 6, 1, 6, 1, 20, 1)), gf.KSW_FIELDS)
    assert np.array_equal(got, d["work_want"])
    g = gf.batch(d, "gate_")
    for a, bb, minsc in G.RESCUE_SCORINGS:
        got = gf.as2d(gpu_ctx.align2_arrays(g, bwg.ksw_params(5, G.scmat(a, bb), 6, 1, 6, 1, minsc, a)),
                      gf.KSW_FIELDS)
        assert np.array_equal(got, d["gate_want_%d_%d_%d" % (a, bb, minsc)]), a
