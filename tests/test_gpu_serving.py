"""GPU tests of the serving loop: results alone == results in the mix, reference-window targets,
HBM-budget admission."""
import sys
import threading

import numpy as np
import pytest

import bwg
import oracle_binding as ob

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


def _params():
    mat = bwg.default_mat()
    return (bwg.ext_params(5, mat, 6, 1, 6, 1, 100, 5, 100), bwg.ksw_params(5, mat, 6, 1, 6, 1, 19, 1), mat)


def test_server_mix_equals_alone(gpu_ctx):
    """Chunks submitted concurrently from several threads, extension and rescue, arena targets and
    genome-window targets, merged into shared epochs: every chunk's results equal the chunk run
    alone, and a sample equals the oracle."""
    _say("server_mix_equals_alone")
    ep, kp, mat = _params()
    gpu_ctx.set_form(bwg.FORM_BATCHED)
    srv = bwg.Server(gpu_ctx, hbm_budget_bytes=8 << 30, slots=3)
    srv.set_reference(bwg.genome(7))
    chunks = []
    rng = np.random.default_rng(1)
    for c in range(24):
        kind = bwg.WORKLOAD_EXT_TAIL if c % 4 else bwg.WORKLOAD_RESCUE
        n = int(rng.integers(200, 6000))
        batch, win = bwg.generate_windows(kind, n, seed=100 + c, genome_seed=7)
        params = ep if kind == bwg.WORKLOAD_EXT_TAIL else kp
        alone = (gpu_ctx.extend_arrays(batch, params) if params is ep else gpu_ctx.align2_arrays(batch, params))
        chunks.append((batch, win if c % 3 == 0 else None, params, alone))
    results = [None] * len(chunks)

    def producer(ids):
        tickets = [(i, srv.submit(chunks[i][0], chunks[i][2], chunks[i][1])) for i in ids]
        for i, tk in tickets:
            results[i] = srv.wait(tk)

    th = [threading.Thread(target=producer, args=(list(range(p, len(chunks), 4)),)) for p in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for i, (batch, win, params, alone) in enumerate(chunks):
        assert results[i] is not None and np.array_equal(results[i], alone), "chunk %d" % i
    batch, win, params, alone = chunks[1]
    assert np.array_equal(alone, ob.extend_batch(batch, mat, 6, 1, 6, 1, 100, 5, 100))
    batch, win, params, alone = chunks[0]
    assert np.array_equal(alone, ob.align2_batch(batch, mat, 6, 1, 6, 1, 19, 1))
    cnt = srv.counters()
    assert cnt["chunks"] == len(chunks) and cnt["epochs"] >= 1
    srv.close()


def test_server_budget_admission(gpu_ctx):
    ep, kp, mat = _params()
    batch = bwg.generate(bwg.WORKLOAD_EXT_TAIL, 3000, seed=3)
    srv = bwg.Server(gpu_ctx, hbm_budget_bytes=1 << 20, slots=2)  # smaller than one chunk
    with pytest.raises(bwg.BwgError) as e:
        srv.submit(batch, ep)
    assert e.value.code == bwg.BWG_E2BIG
    srv.close()
    # a budget that admits one chunk at a time still completes every chunk, in any order
    one = bwg.generate(bwg.WORKLOAD_EXT_TAIL, 500, seed=4)
    srv = bwg.Server(gpu_ctx, hbm_budget_bytes=48 << 20, slots=3)
    tks = [srv.submit(one, ep) for _ in range(6)]
    want = gpu_ctx.extend_arrays(one, ep)
    for tk in tks:
        assert np.array_equal(srv.wait(tk), want)
    assert srv.counters()["peak_device_bytes"] <= 48 << 20
    srv.close()


def test_server_rejects_bad_windows(gpu_ctx):
    ep, kp, mat = _params()
    batch, win = bwg.generate_windows(bwg.WORKLOAD_EXT_TAIL, 100, seed=2, genome_seed=2)
    srv = bwg.Server(gpu_ctx, hbm_budget_bytes=1 << 30, slots=1)
    with pytest.raises(bwg.BwgError):  # no reference uploaded yet
        srv.submit(batch, ep, win)
    srv.set_reference(bwg.genome(2))
    bad = win.copy()
    bad[0] = bwg.TARGET_FROM_REF | (len(bwg.genome(2)) - 5)
    with pytest.raises(bwg.BwgError):
        srv.submit(batch, ep, bad)
    assert np.array_equal(srv.wait(srv.submit(batch, ep, win)), gpu_ctx.extend_arrays(batch, ep))
    srv.close()


def test_serve_driver_poisson_load_bit_exact(gpu_ctx):
    """The C++ load driver: 8 Poisson producers at a rate that overlaps chunks on the device;
    every chunk compared with its alone result."""
    _say("serve_driver")
    srv = bwg.Server(gpu_ctx, hbm_budget_bytes=64 << 30, slots=3)
    r = bwg.serve_run(gpu_ctx, srv, rate_jobs_s=20e6, producers=8, duration_s=2.0, pool_chunks=16, verify=True)
    srv.close()
    assert r["chunks"] > 10 and r["mismatched_chunks"] == 0 and r["failed_chunks"] == 0
