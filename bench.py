#!/usr/bin/env python3
"""Benchmark of the Smith-Waterman hot path on one MI355X.

Headline workload: 1M long-tail extension jobs (reads of 100/150/250 bp, query lengths 70% < 30,
25% 30-150, 5% 150-250, target = query + 100, band 100, h0 U[19,250]), run through the batched
form with the batch resident in HBM.  A step is one pass of the extension kernels over the whole
batch.  Alongside it the same batch is timed in the plain form, the 200k-job mate-rescue workload
is timed in both forms, and the CPU oracle is timed on the host cores.

Prints one JSON line.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "bwa-mem4_amd"))

import numpy as np  # noqa: E402

import bwg  # noqa: E402

METRIC = "extension GCUPS + jobs/s, batched vs plain GPU (bit-exact); p50/p99 chunk latency"
K_EXT = 16   # integer ops per extension cell (crates/bwa-extend/src/sw.rs:377-412)
K_RESC = 13  # integer ops per rescue cell (crates/bwa-extend/src/sw.rs:993-1029)
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12  # int32 lane-ops/s on 256 CUs at 2.4 GHz
HBM_PEAK_GBPS = 8000.0


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def time_steps(ctx, staged, steps, warmup):
    for _ in range(warmup):
        staged.run()
    kern_ns, main_ns, launches = [], [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        staged.run()
        kern_ns.append(ctx.stat(bwg.STAT_LAST_KERNEL_NS))
        main_ns.append(ctx.stat(bwg.STAT_LAST_MAIN_NS))
        launches.append(ctx.stat(bwg.STAT_LAST_MAIN_LAUNCHES))
    wall = time.perf_counter() - t0
    return wall / steps, float(np.mean(kern_ns)) * 1e-9, float(np.mean(main_ns)) * 1e-9, float(np.mean(launches))


def cpu_baseline(ext_batch, mat, n_threads):
    """The oracle on the host cores over the whole extension batch, with the same algorithm the
    batched GPU form runs: the ungapped-diagonal fast path, then ksw_extend2 for the declined jobs
    (one pthread per core, static partition).  Returns the results, the time, the workload's
    evaluated-cell count (the GCUPS denominator) and the cells of the declined jobs (the cells the
    DP kernels evaluate)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_binding as ob
    t0 = time.perf_counter()
    out, hit = ob.ungapped_batch(ext_batch, mat, 6, 1, 6, 1, 100)
    idx = np.nonzero(~hit)[0]
    rest = bwg.JobBatch(ext_batch.arena, ext_batch.q_off[idx], ext_batch.q_len[idx], ext_batch.t_off[idx],
                        ext_batch.t_len[idx], ext_batch.h0[idx])
    out[idx] = ob.extend_batch(rest, mat, 6, 1, 6, 1, 100, 5, 100, nthreads=n_threads)
    dt = time.perf_counter() - t0
    full, job_cells = ob.extend_batch_job_cells(ext_batch, mat, 6, 1, 6, 1, 100, 5, 100, nthreads=n_threads)
    assert np.array_equal(full, out)
    return out, int(job_cells.sum()), dt, int(job_cells[idx].sum()), int(hit.sum())


def traffic_from_profile(jobs):
    """HBM bytes per batch of the long-tail extension workload, from the committed rocprofv3
    counter passes (FETCH_SIZE and WRITE_SIZE, each scaled by the copy-kernel calibration) of the
    same workload; None when that summary is absent or was taken at another batch size."""
    path = os.path.join(ROOT, "profiles", "ext_batched_longtail_1M_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("jobs") != jobs:
        return None
    return t.get("bytes_per_batch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--jobs", type=int, default=1_000_000)
    ap.add_argument("--rescue-jobs", type=int, default=200_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--forms", default="batched,plain", help="comma list of forms to time")
    ap.add_argument("--only", default="ext,small,resc,side,matesw,global,e2e",
                    help="comma list of workloads to time")
    ap.add_argument("--global-jobs", type=int, default=200_000,
                    help="region-to-CIGAR jobs of the banded global alignment workload")
    ap.add_argument("--matesw-anchors", type=int, default=200_000,
                    help="anchors of the mate-rescue round (4 orientation jobs each)")
    ap.add_argument("--serve-rates", default="5e4,2e5,2e6,1e7,3e7,6e7",
                    help="offered loads (jobs/s) of the serving sweep; empty to skip")
    ap.add_argument("--serve-seconds", type=float, default=3.0,
                    help="seconds per offered load (longer at low loads, for at least ~40 chunks)")
    ap.add_argument("--small-jobs", type=int, default=100_000,
                    help="jobs of the small plain-extension workload (0 to skip)")
    ap.add_argument("--serial-bins", action="store_true",
                    help="run a batch's bins one after another on one stream (for counter runs)")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="also run one device-to-device copy of this size (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this path has no multi-GPU step (jobs shard across GPUs "
              "with no exchange); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    mat = bwg.default_mat()
    ext_p = bwg.ext_params(5, mat, 6, 1, 6, 1, 100, 5, 100)
    ksw_p = bwg.ksw_params(5, mat, 6, 1, 6, 1, 20, 1)
    log("generating %d long-tail extension jobs, %d small extension jobs and %d rescue jobs"
        % (args.jobs, args.small_jobs, args.rescue_jobs))
    ext = bwg.generate(bwg.WORKLOAD_EXT_TAIL, args.jobs, seed=42)
    resc = bwg.generate(bwg.WORKLOAD_RESCUE, args.rescue_jobs, seed=42)
    small = bwg.generate(bwg.WORKLOAD_EXT_SMALL, args.small_jobs, seed=42) if args.small_jobs else None

    ctx = bwg.Context(0)
    if args.serial_bins:
        ctx.set_option(bwg.OPT_CONCURRENT_BINS, 0)
    copy_gbs = None
    if args.calib_copy_mb:
        copy_gbs = ctx.hbm_copy(args.calib_copy_mb << 20, reps=1)
        log("calibration copy: hbm_copy_kernel, %d MiB read + %d MiB written per launch, %.0f GB/s"
            % (args.calib_copy_mb, args.calib_copy_mb, copy_gbs))
    res = {}
    outs = {}
    forms = [f for f in ((bwg.FORM_BATCHED, "batched"), (bwg.FORM_PLAIN, "plain")) if f[1] in args.forms.split(",")]
    only = args.only.split(",")
    for form, fname in forms:
        try:
            ctx.set_form(form)
            if "ext" not in only:
                raise KeyError
            log("extension %s: staging" % fname)
            st = ctx.stage(ext, ext_p)
            st.run()  # first run builds the form's plan and buffers
            per, kern, main, nl = time_steps(ctx, st, args.steps if form == bwg.FORM_BATCHED else max(1, args.steps // 3), args.warmup)
            outs["ext_" + fname] = st.results()
            hits = st.ungapped_hits()
            st.close()
            res["ext_" + fname] = dict(step_s=per, kernel_s=kern, main_s=main, main_launches=nl, ungapped_hits=hits)
            log("extension %s: %.3f ms/step (kernels %.3f ms), %d jobs resolved by the ungapped fast path"
                % (fname, per * 1e3, kern * 1e3, hits))
            if form == bwg.FORM_BATCHED:
                # the same batch with every job through the banded DP, for comparison
                ctx.set_option(bwg.OPT_UNGAPPED, 0)
                st = ctx.stage(ext, ext_p)
                st.run()
                per0, kern0, _, _ = time_steps(ctx, st, max(1, args.steps // 2), 1)
                assert np.array_equal(st.results(), outs["ext_batched"]), "fast path changes results"
                st.close()
                ctx.set_option(bwg.OPT_UNGAPPED, 1)
                res["ext_batched_dp_only"] = dict(step_s=per0, kernel_s=kern0)
                log("extension batched, fast path off: %.3f ms/step (kernels %.3f ms)" % (per0 * 1e3, kern0 * 1e3))
        except KeyError:
            pass
        except bwg.BwgError as e:
            log("form %s unavailable: %s" % (fname, e))
        try:
            if "small" not in only or small is None:
                raise KeyError
            st = ctx.stage(small, ext_p)
            st.run()
            per, kern, main, nl = time_steps(ctx, st, args.steps, args.warmup)
            outs["small_" + fname] = st.results()
            st.close()
            res["small_" + fname] = dict(step_s=per, kernel_s=kern, main_s=main, main_launches=nl)
            log("small extension %s: %.3f ms/step (kernels %.3f ms)" % (fname, per * 1e3, kern * 1e3))
        except KeyError:
            pass
        try:
            if "resc" not in only:
                raise KeyError
            log("rescue %s: staging" % fname)
            st = ctx.stage(resc, ksw_p)
            st.run()
            per, kern, main, nl = time_steps(ctx, st, max(1, args.steps // 2) if form == bwg.FORM_BATCHED else 2, 1)
            outs["resc_" + fname] = st.results()
            st.close()
            res["resc_" + fname] = dict(step_s=per, kernel_s=kern, main_s=main, main_launches=nl)
            log("rescue %s: %.3f ms/step (kernels %.3f ms)" % (fname, per * 1e3, kern * 1e3))
        except KeyError:
            pass
        except bwg.BwgError as e:
            log("form %s unavailable: %s" % (fname, e))
    side = {}
    if "side" in only:
        # one side of the seed extensions with bwa's band doubling, all rounds on the device: the
        # long-tail batch at w = 100 (as above) and at w = 5, where a few percent of the jobs fail
        # the acceptance test and run again at w = 10; left (-1) and right (h0) jobs alternate
        ctx.set_form(bwg.FORM_BATCHED)
        prev = np.where(np.arange(args.jobs) % 2 == 0, -1, ext.h0).astype(np.int32)
        for w in (100, 5):
            p = bwg.ext_params(5, mat, 6, 1, 6, 1, w, 5, 100)
            log("side batch w=%d: staging" % w)
            st = ctx.stage_side(ext, p, prev, 2)
            st.run()
            per, kern, _, _ = time_steps(ctx, st, max(1, args.steps // 2), 1)
            bands = st.side_bands()
            if w == 100 and "ext_batched" in outs:
                assert np.array_equal(st.results(), outs["ext_batched"]), "side batch differs from the plain batch"
            st.close()
            side[w] = dict(step_s=per, kernel_s=kern, retried=int((bands != w).sum()))
            log("side batch w=%d: %.3f ms/step (kernels %.3f ms), %d jobs retried at w=%d"
                % (w, per * 1e3, kern * 1e3, side[w]["retried"], 2 * w))
    matesw = None
    if "matesw" in only and args.matesw_anchors:
        # one mate-rescue round with device-resident windows: a 64 Mbp synthetic genome in 8
        # contigs, one anchor per read pair (150 bp reads, insert N(550,150), 30% unplaced mates)
        ctx.set_form(bwg.FORM_BATCHED)
        log("mate rescue: generating %d anchors on a 64 Mbp genome" % args.matesw_anchors)
        rnd = bwg.generate_rescue_round(args.matesw_anchors, genome_len=64_000_000, n_contigs=8, seed=43)
        ctx.set_reference(rnd.fwd, rnd.contigs)
        msw_p = bwg.ksw_params(5, mat, 6, 1, 6, 1, rnd.min_seed_len, 1)  # minsc = min_seed_len * a
        st = ctx.stage_matesw(rnd, msw_p)
        st.run()
        per, kern, _, _ = time_steps(ctx, st, max(1, args.steps // 2), 1)
        win, out = st.windows(), st.results()
        st.close()
        valid = (win["flags"] & bwg.WIN_VALID) != 0
        cells = float(np.sum((win["re"] - win["rb"])[valid].astype(np.int64) *
                             rnd.anchors["l_ms"].repeat(4)[valid].astype(np.int64)))
        matesw = dict(step_s=per, kernel_s=kern, jobs=int(valid.sum()), cells=cells)
        log("mate rescue: %.3f ms/step (kernels %.3f ms), %d windows aligned, %.0f G cells/s"
            % (per * 1e3, kern * 1e3, matesw["jobs"], cells / per / 1e9))
        if not args.no_check:
            sys.path.insert(0, os.path.join(ROOT, "tests"))
            import oracle_binding as ob
            sub_n = min(10_000, args.matesw_anchors)
            part = bwg.RescueRound(rnd.fwd, rnd.contigs, rnd.pes, rnd.anchors[:sub_n], rnd.ma_rb[:rnd.n_ma],
                                   rnd.arena, rnd.min_seed_len)
            want_w, want = ob.matesw_batch(part, mat, 6, 1, 6, 1, rnd.min_seed_len, 1, nthreads=args.cpu_threads)
            ok = np.array_equal(want_w, win[:4 * sub_n]) and np.array_equal(want, out[:4 * sub_n])
            log("mate rescue: first %d anchors bit-exact vs oracle: %s" % (sub_n, ok))
            if not ok:
                sys.exit(1)
    glb = None
    if "global" in only and args.global_jobs:
        # region-to-CIGAR: reads of 100/150/250 bp against their reference span (1% substitutions,
        # 30% of the regions with indels), each with the band the reference's CIGAR step gives
        ctx.set_form(bwg.FORM_BATCHED)
        log("global alignment: generating %d region jobs" % args.global_jobs)
        gb, gw = bwg.generate_global_jobs(args.global_jobs, seed=44)
        gp = bwg.global_params(5, mat, 6, 1, 6, 1)
        st = ctx.stage_global(gb, gp, gw)
        st.run()
        per, kern, _, _ = time_steps(ctx, st, max(1, args.steps // 2), 1)
        gscore = st.results()
        gn, gc = st.cigars()
        st.close()
        qg, tg = gb.q_len.astype(np.int64), gb.t_len.astype(np.int64)
        gcells = float(np.sum(np.minimum(qg, 2 * gw.astype(np.int64) + 1) * tg))  # upper bound: full band rows
        glb = dict(step_s=per, kernel_s=kern, cells=gcells)
        log("global alignment: %.3f ms/step (kernels %.3f ms), %.3g jobs/s" % (per * 1e3, kern * 1e3,
                                                                                 args.global_jobs / per))
        if not args.no_check:
            sys.path.insert(0, os.path.join(ROOT, "tests"))
            import oracle_binding as ob
            ws, wn, wc = ob.global_batch(gb, gw, mat, 6, 1, 6, 1, nthreads=args.cpu_threads)
            ok = np.array_equal(ws, gscore) and np.array_equal(wn.astype(np.uint32), gn) and np.array_equal(wc, gc)
            log("global alignment: bit-exact vs oracle: %s" % ok)
            if not ok:
                sys.exit(1)
    e2e = None
    if "e2e" in only and "ext" in only:
        # PCIe-inclusive: the synchronous entry point from host buffers (stage, run, copy back)
        ctx.set_form(bwg.FORM_BATCHED)
        ctx.extend_arrays(ext, ext_p)
        t0 = time.perf_counter()
        reps = 3
        for _ in range(reps):
            ctx.extend_arrays(ext, ext_p)
        e2e = args.jobs * reps / (time.perf_counter() - t0)
        log("extension from host buffers (PCIe-inclusive): %.3g jobs/s" % e2e)
    serving = []
    rates = [float(x) for x in args.serve_rates.split(",") if x.strip()]
    if rates:
        ctx.set_form(bwg.FORM_BATCHED)
        srv = bwg.Server(ctx, hbm_budget_bytes=64 << 30, slots=3)
        for rate in rates:
            # chunks average ~15k jobs: run long enough at low loads for ~40 of them
            secs = max(args.serve_seconds, min(15.0, 40 * 15000.0 / rate))
            r = bwg.serve_run(ctx, srv, rate, producers=8, duration_s=secs, seed=42,
                              use_ref=True, pool_chunks=64, verify=True)
            r["offered_jobs_per_s"] = rate
            serving.append(r)
            log("serving at %.3g jobs/s: p50 %.2f ms p99 %.2f ms, completed %.3g jobs/s, %d chunks, "
                "%d mismatched" % (rate, r["p50_ms"], r["p99_ms"], r["jobs_per_s"], r["chunks"],
                                   r["mismatched_chunks"]))
            if r["mismatched_chunks"] or r["failed_chunks"]:
                log("serving results differ from the chunks run alone")
                sys.exit(1)
        srv.close()
    head = "ext_batched" if "ext_batched" in res else "ext_plain"
    if head not in res:
        log("no extension timing requested; nothing to report")
        return
    if "ext_batched" in outs and "ext_plain" in outs:
        assert np.array_equal(outs["ext_batched"], outs["ext_plain"]), "forms disagree"
    if "small_batched" in outs and "small_plain" in outs:
        assert np.array_equal(outs["small_batched"], outs["small_plain"]), "forms disagree"

    cpu = None
    cells = None
    small_cells = None
    dp_cells = cpu_hits = None
    if not args.no_cpu:
        log("cpu oracle over the extension batch on %d threads" % args.cpu_threads)
        cpu_out, cells, cpu_dt, dp_cells, cpu_hits = cpu_baseline(ext, mat, args.cpu_threads)
        if not args.no_check and head in outs:
            ok = np.array_equal(cpu_out, outs[head])
            log("bit-exact vs oracle: %s" % ok)
            if not ok:
                bad = np.nonzero(cpu_out != outs[head])[0]
                log("mismatches: %d, first %d" % (bad.size, bad[0]))
                sys.exit(1)
        cpu = {"value": args.jobs / cpu_dt, "unit": "jobs/s", "cores": args.cpu_threads,
               "kind": "port", "gcups": cells / cpu_dt / 1e9,
               "sample": "all %d long-tail extension jobs: the ungapped-diagonal fast path, then the "
                         "scalar C oracle's ksw_extend2 for the declined jobs, static partition" % args.jobs}
        small_key = "small_batched" if "small_batched" in outs else "small_plain"
        if small_key in outs:
            s_out, small_cells, _, _, _ = cpu_baseline(small, mat, args.cpu_threads)
            ok = np.array_equal(s_out, outs[small_key])
            log("small workload bit-exact vs oracle: %s" % ok)
            if not ok:
                sys.exit(1)

    h = res[head]
    jobs_s = args.jobs / h["step_s"]
    line = {
        "metric": METRIC, "value": jobs_s, "unit": "jobs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": h["step_s"] * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": "batched uneven extension: %d ksw_extend2 jobs, reads 100/150/250bp, "
                               "qlen 70%%<30/25%% 30-150/5%% 150-250, tlen=qlen+100, w=100, zdrop=100, "
                               "end_bonus=5, h0 U[19,250], seed 42" % args.jobs,
                   "form": head.split("_")[1]},
    }
    if cells:
        line["gcups"] = cells / h["step_s"] / 1e9
        line["gcups_kernel"] = cells / h["kernel_s"] / 1e9
        line["cells"] = cells
        # the DP kernels evaluate the cells of the jobs the ungapped fast path declines
        if h.get("ungapped_hits"):
            assert h["ungapped_hits"] == cpu_hits, "device and oracle fast paths disagree"
            line["ungapped_hits"] = h["ungapped_hits"]
            line["dp_cells"] = dp_cells
        ops = (dp_cells if h.get("ungapped_hits") else cells) * K_EXT / h["kernel_s"] / 1e12
        # algorithmic HBM bytes: query and target bytes, descriptor, order entry and result per job
        alg_bytes = float(np.sum(ext.q_len.astype(np.int64) + ext.t_len)) + args.jobs * (32 + 4 + 24)
        line["roofline"] = {"bound": "valu", "achieved": ops, "peak": VALU_PEAK_TOPS,
                            "unit": "Tops/s", "frac": ops / VALU_PEAK_TOPS,
                            "traffic": traffic_from_profile(args.jobs),
                            "hbm": {"achieved": alg_bytes / h["kernel_s"] / 1e9, "peak": HBM_PEAK_GBPS,
                                    "unit": "GB/s", "frac": alg_bytes / h["kernel_s"] / 1e9 / HBM_PEAK_GBPS,
                                    "alg_bytes": alg_bytes}}
    if "ext_batched_dp_only" in res:
        line["batched_dp_only_jobs_per_s"] = args.jobs / res["ext_batched_dp_only"]["step_s"]
        if cells:
            line["batched_dp_only_roofline_frac"] = cells * K_EXT / res["ext_batched_dp_only"]["kernel_s"] / 1e12 / VALU_PEAK_TOPS
    if "ext_plain" in res:
        line["plain_jobs_per_s"] = args.jobs / res["ext_plain"]["step_s"]
        line["speedup_vs_plain"] = jobs_s / line["plain_jobs_per_s"]
    if e2e:
        line["e2e_jobs_per_s"] = e2e
    if matesw:
        line["matesw_anchors_per_s"] = args.matesw_anchors / matesw["step_s"]
        line["matesw_jobs_per_s"] = matesw["jobs"] / matesw["step_s"]
        line["matesw_gcell_s"] = matesw["cells"] / matesw["step_s"] / 1e9
    if glb:
        line["global_jobs_per_s"] = args.global_jobs / glb["step_s"]
        line["global_band_gcell_s"] = glb["cells"] / glb["step_s"] / 1e9
    for w, r in side.items():
        line["side_w%d_jobs_per_s" % w] = args.jobs / r["step_s"]
        line["side_w%d_retried" % w] = r["retried"]
    for f in ("batched", "plain"):
        if "small_" + f in res:
            r = res["small_" + f]
            line["small_%s_jobs_per_s" % f] = args.small_jobs / r["step_s"]
            if cpu and small_cells:
                line["small_%s_gcups" % f] = small_cells / r["step_s"] / 1e9
    for f in ("batched", "plain"):
        if "resc_" + f in res:
            r = res["resc_" + f]
            line["rescue_%s_jobs_per_s" % f] = args.rescue_jobs / r["step_s"]
            line["rescue_%s_gcell_s" % f] = float(np.sum(resc.q_len.astype(np.int64) * resc.t_len)) / r["step_s"] / 1e9
    if serving:
        at_cfg = min(serving, key=lambda r: abs(r["offered_jobs_per_s"] - 2e5))
        line["p50_chunk_latency_ms"] = at_cfg["p50_ms"]
        line["p99_chunk_latency_ms"] = at_cfg["p99_ms"]
        line["serving"] = [{k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()}
                           for r in serving]
    rep = ctx.tune_report()
    if rep:
        line["tuning"] = rep.splitlines()
    if cpu:
        line["cpu_baseline"] = cpu
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
