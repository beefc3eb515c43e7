#!/usr/bin/env python3
"""Benchmark of the Smith-Waterman hot path on one MI355X.

Headline workload: 1M long-tail extension jobs (reads of 100/150/250 bp, query lengths 70% < 30,
25% 30-150, 5% 150-250, target = query + 100, band 100, h0 U[19,250]), run through the batched
form with the batch resident in HBM.  A step is one pass of the extension kernels over the whole
batch.  Alongside it:

- the same batch in the plain form, and per-bin device times with the bins run one after another
  (each bin's own fraction of the integer-VALU roofline);
- the 200k-job mate-rescue workload (minsc 20) in both forms, checked bit-exact against the oracle,
  with its own roofline, executed-cell count and CPU baseline;
- the small extension workload, band retry, a mate-rescue round, global alignment;
- the reference's call shapes: 16 host threads making 2000-job extension calls and 100-job rescue
  calls on one context, directly and through the server;
- the synchronous call from host buffers (PCIe included) with its phases;
- the Poisson serving sweep with 8 producers;
- the CPU oracle on the host cores.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "bwa-mem4_amd"))

import numpy as np  # noqa: E402

import bwg  # noqa: E402

METRIC = "extension GCUPS + jobs/s, batched vs plain GPU (bit-exact); p50/p99 chunk latency"
K_EXT = 16   # integer ops per extension cell (crates/bwa-extend/src/sw.rs:377-412)
K_RESC = 13  # integer ops per rescue cell (crates/bwa-extend/src/sw.rs:993-1029)
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12  # int32 lane-ops/s on 256 CUs at 2.4 GHz
VALU_PK16_PEAK_TOPS = 2 * VALU_PEAK_TOPS  # packed 16-bit lane-ops/s
HBM_PEAK_GBPS = 8000.0
EXT_P = (6, 1, 6, 1, 100, 5, 100)  # o_del e_del o_ins e_ins w end_bonus zdrop
RESC_P = (6, 1, 6, 1, 20, 1)      # o_del e_del o_ins e_ins minsc max_sc (config 2: minsc 20)


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def oracle():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_binding as ob
    return ob


def host_info():
    """CPU model, online CPUs, CPUs this process may run on, and the worker threads the CPU legs
    use: the CPUs available to the process, at most 16 (a one-GPU box's CPU share)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    online = os.cpu_count() or 1
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = online
    return {"cpu_model": model, "cpus_online": online, "cpus_available": avail, "threads": min(16, avail)}


SEGMENTS = []  # labels of the trace segments the marker copies delimit, in order


def mark(ctx, label, **info):
    """A 1 MiB device copy (hbm_copy_kernel) as a marker in a kernel trace: the dispatches between
    two markers belong to the segment `label` names (--segments-out, --summarize-trace); `info`
    goes into that segment's entry."""
    ctx.hbm_copy(1 << 20, reps=1)
    SEGMENTS.append(dict(label=label, **info))


def segment(label):
    return next(s for s in SEGMENTS if s["label"] == label)


def time_steps(ctx, staged, steps, warmup):
    for _ in range(warmup):
        staged.run()
    kern_ns = []
    t0 = time.perf_counter()
    for _ in range(steps):
        staged.run()
        kern_ns.append(ctx.stat(bwg.STAT_LAST_KERNEL_NS))
    wall = time.perf_counter() - t0
    return wall / steps, float(np.mean(kern_ns)) * 1e-9


def serial_bin_times(ctx, staged, reps=3):
    """Each launch's device time with the batch's bins run one after another on one stream
    (HIP events around every launch), averaged over `reps` runs; and the report of the last."""
    ctx.set_option(bwg.OPT_CONCURRENT_BINS, 0)
    ctx.set_option(bwg.OPT_ITEM_TIMING, 1)
    try:
        acc = None
        for _ in range(reps):
            staged.run()
            rep = staged.report()
            us = np.array([r["us"] for r in rep])
            acc = us if acc is None else acc + us
        for r, t in zip(rep, acc / reps):
            r["us"] = float(t)
        total_s = ctx.stat(bwg.STAT_LAST_KERNEL_NS) * 1e-9
    finally:
        ctx.set_option(bwg.OPT_ITEM_TIMING, 0)
        ctx.set_option(bwg.OPT_CONCURRENT_BINS, 1)
    return rep, total_s


def ext_bin_table(rep, bins, job_cells, dp_mask):
    """Per extension DP bin: jobs, jobs the DP ran (declined by the ungapped pre-pass), their
    evaluated cells, serial device time and fraction of the int32 VALU roofline."""
    out = []
    for r in rep:
        sel = bins == r["index"]
        if not sel.any() and r["kernel"] != "ext_ungapped":
            continue
        row = {"kernel": r["kernel"], "jobs": int(sel.sum()), "us": round(r["us"], 2)}
        if r["kernel"] == "ext_ungapped":
            row["jobs"] = r["jobs"]
        else:
            dsel = sel & dp_mask
            cells = int(job_cells[dsel].sum())
            row.update(dp_jobs=int(dsel.sum()), dp_cells=cells)
            row.update(ops=cells * K_EXT, peak_tops=VALU_PEAK_TOPS)
            if r["us"] > 0:
                ops = cells * K_EXT / (r["us"] * 1e-6) / 1e12
                row["frac"] = round(ops / VALU_PEAK_TOPS, 4)
        out.append(row)
    return out


def resc_bin_table(rep, bins, batch, res, max_sc=1):
    """Per rescue launch: forward bins with their nominal (qlen x tlen) and padded cells and the
    fraction of the roofline of their arithmetic (packed 16-bit or int32); the tracker; the
    reverse bins with the cells of the reversed prefixes up to the forward score."""
    q = batch.q_len.astype(np.int64)
    t = batch.t_len.astype(np.int64)
    lanes = np.where(q * max_sc < 250, 16, 8)
    qpad = (q + lanes - 1) // lanes * lanes
    rq = np.minimum(res["qe"].astype(np.int64) + 1, q)
    rl = np.where(rq * max_sc < 250, 16, 8)
    rpad = (rq + rl - 1) // rl * rl
    rrows = np.where(res["tb"] >= 0, res["te"].astype(np.int64) - res["tb"] + 1, res["te"].astype(np.int64) + 1)
    ran_rev = (res["score"] >= RESC_P[4]) & (res["qe"] >= 0)
    fwd_of = {}
    out = []
    for r in rep:
        row = {"kernel": r["kernel"], "jobs": r["jobs"], "us": round(r["us"], 2)}
        sel = bins == r["index"]
        packed = r["kernel"].startswith("ksw_pk")
        peak = VALU_PK16_PEAK_TOPS if packed else VALU_PEAK_TOPS
        if sel.any():
            fwd_of[r["kernel"].split(">")[0]] = sel
            row["cells"] = int((q[sel] * t[sel]).sum())
            row["padded_cells"] = int((qpad[sel] * t[sel]).sum())
            row.update(ops=row["padded_cells"] * K_RESC, peak_tops=peak)
            if r["us"] > 0:
                row["frac"] = round(row["padded_cells"] * K_RESC / (r["us"] * 1e-6) / 1e12 / peak, 4)
                row["peak_tops"] = peak
        elif "reverse" in r["kernel"]:
            sel = fwd_of.get(r["kernel"].split(",reverse")[0].split(">")[0], np.zeros(len(q), bool)) & ran_rev
            row["cells"] = int((rpad[sel] * rrows[sel]).sum())
            row.update(ops=row["cells"] * K_RESC, peak_tops=peak)
            if r["us"] > 0:
                row["frac"] = round(row["cells"] * K_RESC / (r["us"] * 1e-6) / 1e12 / peak, 4)
                row["peak_tops"] = peak
        out.append(row)
    rev_cells = int((rpad[ran_rev] * rrows[ran_rev]).sum())
    return out, int((qpad * t).sum()), rev_cells


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


def _short_kernel(name):
    """A trace's demangled kernel name without return type, namespaces and argument list."""
    name = name.replace("(anonymous namespace)::", "").split("(", 1)[0]
    if name.startswith("void "):
        name = name[5:]
    head, sep, tmpl = name.partition("<")
    return head.rsplit("::", 1)[-1] + sep + tmpl


def _trace_name(bin_name):
    """The trace name of the kernel a bench bin label (the library's mapping name) launches."""
    import re
    m = re.match(r"(ext_group8|ext_group|ext_thread|ksw_pk8|ksw_pk|ksw_group)<(\d+),i(\d+)(,reverse)?>$", bin_name)
    if not m:
        return bin_name + "_kernel"
    fam, k, width, rev = m.group(1), int(m.group(2)), int(m.group(3)), bool(m.group(4))
    narrow = "true" if width == 16 else "false"
    if fam in ("ext_group", "ext_group8"):
        return "ext_group_kernel<%d, %s, %d>" % (k, narrow, 8 if fam == "ext_group8" else 16)
    if fam == "ext_thread":
        return "ext_thread_kernel<%d, %s>" % (k, narrow)
    if fam in ("ksw_pk", "ksw_pk8"):  # template: column pairs per lane, direction, group lanes
        return "ksw_pk_kernel<%d, %s, %d>" % (k // 2, "true" if rev else "false", 8 if fam == "ksw_pk8" else 16)
    return "%s_kernel<%d, %s>" % (fam, k, "true" if rev else "false")


def _split_at_markers(rows, what, with_markers=False):
    """Dispatch rows (kernel name at [2], grid at [3], in dispatch order) split at the marker copies:
    the rows before the first marker, then those after each marker (and, with_markers, each
    marker's own rows)."""
    copies = [r for r in rows if r[2] == "hbm_copy_kernel"]
    if not copies:
        raise SystemExit("no marker copies in %s" % what)
    grid = copies[-1][3]  # the run ends with a marker: its grid is the markers' grid
    # a marker is the fill of the copy's source buffer and then the copies (a warm-up and the timed one)
    is_copy = [r[2] == "hbm_copy_kernel" and r[3] == grid for r in rows]
    starts = {i for i in range(len(rows) - 1) if rows[i][2].startswith("__amd_rocclr_fillBuffer") and is_copy[i + 1]}
    in_marker = starts | {i for i, c in enumerate(is_copy) if c}
    pieces, cur, marks = [], [], []
    for i, r in enumerate(rows):
        if i in starts:
            pieces.append(cur)
            cur = []
            marks.append([])
        if i not in in_marker:
            cur.append(r)
        elif marks:
            marks[-1].append(r)
    pieces.append(cur)
    return (pieces, marks) if with_markers else pieces


def _launch_sequences(piece):
    """A segment's dispatches split into launch sequences: each starts with the zero_u32_kernel
    that clears the queue heads, or with the extension's ungapped pre-pass, which clears them
    itself."""
    runs, cur = [], None
    for r in piece:
        if r[2].startswith("zero_u32_kernel") or r[2].startswith("ext_ungapped_kernel"):
            cur = []
            runs.append(cur)
        if cur is not None:
            cur.append(r)
    return runs


def summarize_trace(trace_csv, segments_json):
    """Markdown summary of a rocprofv3 kernel trace (--kernel-trace --output-format csv) of a bench
    run made with --segments-out: the dispatches split at the marker copies into the labelled
    segments, per-kernel device times per segment, the device span of every timed step (the first
    dispatch of a step clears the queue counters: the zero_u32_kernel, or the extension's ungapped
    pre-pass) with the roofline fraction that span
    gives, and the serial per-bin device times with each bin's fraction recomputed from the trace
    beside the figure bench.py took with HIP events."""
    import csv
    with open(segments_json) as f:
        meta = json.load(f)
    rows = []
    with open(trace_csv) as f:
        for r in csv.DictReader(f):
            rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), _short_kernel(r["Kernel_Name"]),
                         int(r["Grid_Size_X"])))
    rows.sort()
    pieces = _split_at_markers(rows, trace_csv)
    segs = meta["segments"]
    if len(pieces) != len(segs) + 1:
        raise SystemExit("%d marker runs in the trace, %d segments in %s" % (len(pieces) - 1, len(segs),
                                                                              segments_json))
    out = ["# Kernel trace by bench segment", "",
           "%d dispatches; %d before the first marker (staging, the first run that plans, tunes and "
           "captures, the CPU-side set-up)." % (len(rows), len(pieces[0])), ""]

    for s, piece in zip(segs, pieces[1:]):
        label = s["label"]
        out.append("## %s" % label)
        out.append("")
        if not piece:
            out += ["(no dispatches)", ""]
            continue
        by = {}
        for r in piece:
            by.setdefault(r[2], []).append((r[1] - r[0]) / 1e3)
        out.append("| kernel | calls | total us | avg us | min us | max us |")
        out.append("|---|---:|---:|---:|---:|---:|")
        for k, v in sorted(by.items(), key=lambda kv: -sum(kv[1])):
            out.append("| `%s` | %d | %.1f | %.2f | %.2f | %.2f |" % (k, len(v), sum(v), np.mean(v), min(v), max(v)))
        out.append("")
        runs = _launch_sequences(piece)
        if label.endswith("_timed") and runs:
            w = s.get("warmup", 0)
            spans = [(max(r[1] for r in run) - run[0][0]) / 1e3 for run in runs[w:]]
            busy = [sum(r[1] - r[0] for r in run) / 1e3 for run in runs[w:]]
            out.append("%d launch sequences (%d warm-up, %d timed). Timed steps: device span %.1f us mean "
                       "(min %.1f, max %.1f); summed kernel time %.1f us mean (the bins overlap on "
                       "several streams)." % (len(runs), w, len(spans), np.mean(spans), min(spans), max(spans),
                                              np.mean(busy)))
            rl = s.get("roofline")
            if rl:
                ach = rl["ops"] / (np.mean(spans) * 1e-6) / 1e12
                out.append("")
                out.append("Roofline from the trace: %.4g ops per step / %.1f us = %.3f Tops/s = %.4f of "
                           "%.1f Tops/s (%s); bench.py's HIP-event figure: %.4f." %
                           (rl["ops"], np.mean(spans), ach, ach / rl["peak"], rl["peak"], rl["what"],
                            rl["bench_frac"]))
            out.append("")
        if label.endswith("_serial_bins") and runs:
            table = meta.get("ext_bins" if label.startswith("ext") else "rescue_bins") or []
            out.append("%d serial runs, one launch at a time on one stream; per bin the trace's mean "
                       "duration over the runs, and the roofline fraction it gives:" % len(runs))
            out.append("")
            out.append("| bin | trace kernel | jobs | trace avg us | bench.py event us | ops | frac (trace) | "
                       "frac (bench.py) |")
            out.append("|---|---|---:|---:|---:|---:|---:|---:|")
            for b in table:
                tk = _trace_name(b["kernel"])
                d = [(r[1] - r[0]) / 1e3 for run in runs for r in run if r[2] == tk]
                if not d:
                    out.append("| %s | `%s` | %d | (not in trace) | %.1f | | | |" % (b["kernel"], tk, b["jobs"], b["us"]))
                    continue
                t = float(np.mean(d))
                ops, fr = b.get("ops"), ""
                if ops and b.get("peak_tops"):
                    fr = "%.4f" % (ops / (t * 1e-6) / 1e12 / b["peak_tops"])
                out.append("| %s | `%s` | %d | %.1f | %.1f | %s | %s | %s |" %
                           (b["kernel"], tk, b["jobs"], t, b["us"], "%.4g" % ops if ops else "", fr,
                            b.get("frac", "")))
            out.append("")
    print("\n".join(out))


def summarize_counters(paths, segments_json, traffic_out=""):
    """Markdown summary of rocprofv3 --pmc runs (--output-format csv) of one bench command made with
    --segments-out and --calib-copy-mb, one counter set per run: FETCH_SIZE and WRITE_SIZE (KiB per
    dispatch) scaled by what they read for the calibration copy, which moves calib_copy_mb MiB each
    way per launch; per-kernel means over the timed and serial segments (tuning and staging
    excluded); and the extension batch's HBM bytes per launch sequence, written as JSON to
    `traffic_out` when given."""
    import csv
    with open(segments_json) as f:
        meta = json.load(f)
    calib_mb = meta.get("calib_copy_mb") or 0
    runs = []
    for path in paths:
        per = {}
        with open(path) as f:
            for r in csv.DictReader(f):
                did = int(r["Dispatch_Id"])
                grid = int(r.get("Grid_Size") or r.get("Grid_Size_X") or 0)
                e = per.setdefault(did, [_short_kernel(r["Kernel_Name"]), grid, {}])
                e[2][r["Counter_Name"]] = e[2].get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
        runs.append([(did, did, v[0], v[1], v[2]) for did, v in sorted(per.items())])
    # each run on its own (the autotuner may choose other kernels in another run): its segments, the
    # calibration copies before the first marker, and per kernel the counters of every launch in the
    # timed and serial segments
    segs = meta["segments"]
    labels = [s["label"] for s in segs]
    by = {}
    calib = {}
    split = []
    for path, rows in zip(paths, runs):
        pieces, marks = _split_at_markers(rows, path, with_markers=True)
        if len(pieces) != len(segs) + 1:
            raise SystemExit("%d marker runs in %s, %d segments in %s" % (len(pieces) - 1, path, len(segs),
                                                                          segments_json))
        split.append(pieces)
        if segs and segs[0]["label"] == "setup":  # the first copy is the calibration copy
            for r in marks[0]:
                if r[2] == "hbm_copy_kernel":
                    for name, v in r[4].items():
                        calib.setdefault(name, []).append(v)
        for s, piece in zip(segs, pieces[1:]):
            if s["label"].endswith(("_timed", "_serial_bins")):
                for r in piece:
                    for name, v in r[4].items():
                        by.setdefault(r[2], {}).setdefault(name, []).append(v)
    out = ["# Hardware counters by kernel", ""]
    fscale = wscale = None
    if calib_mb and ("FETCH_SIZE" in calib or "WRITE_SIZE" in calib):
        kib = calib_mb * 1024.0
        f, w = calib.get("FETCH_SIZE"), calib.get("WRITE_SIZE")
        fscale = kib / np.mean(f) if f else None
        wscale = kib / np.mean(w) if w else None
        out.append("**Calibration.** `hbm_copy_kernel` moves %d MiB each way per launch; it reads FETCH_SIZE %s KiB "
                   "and WRITE_SIZE %s KiB against %d KiB, so FETCH_SIZE is scaled by %s and WRITE_SIZE by %s "
                   "below." % (calib_mb, "%.0f" % np.mean(f) if f else "-", "%.0f" % np.mean(w) if w else "-",
                               kib, "%.4f" % fscale if fscale else "-", "%.4f" % wscale if wscale else "-"))
        out.append("")

    def mean(cs, key, scale=1.0):
        v = cs.get(key)
        return float(np.mean(v)) * scale if v else None

    def fmt(v, f):
        return f % v if v is not None else ""

    out.append("Per launch, mean over the launches in the timed and serial segments.  fetch / write: scaled "
               "HBM MiB.  VALUBusy and LDSBankConflict: rocprofv3's derived metrics (percent of GPU time "
               "the VALUs issue / LDS is stalled by bank conflicts).  LDS busy: 100 x SQ_LDS_IDX_ACTIVE / (256 "
               "CUs x GPU cycles), the share of cycles the LDS serves indexed accesses, conflicts included; "
               "the CSV's raw GRBM_GUI_ACTIVE is the sum over the 8 XCDs, so GPU cycles are a 1/8 of it (the "
               "derived VALUBusy takes the per-XCD maximum, and equals 100 x SQ_ACTIVE_INST_VALU / (256 x "
               "GRBM_GUI_ACTIVE / 8) here).  VALU util: SQ_THREAD_CYCLES_VALU / (64 x SQ_ACTIVE_INST_VALU), "
               "active lanes per VALU instruction cycle.  LDS conflict / active: SQ_LDS_BANK_CONFLICT / "
               "SQ_LDS_IDX_ACTIVE.")
    out.append("")
    out.append("| kernel | launches | fetch MiB | write MiB | VALU insts | VALUBusy % | LDS busy % | VALU util % | "
               "LDSBankConflict % | LDS conflict / active % |")
    out.append("|---|---:|---:|---:|---:|---:|---:|---:|---:|---:|")
    for k, lst in sorted(by.items(), key=lambda kv: kv[0]):
        n_launch = max(len(v) for v in lst.values())
        fe = mean(lst, "FETCH_SIZE", (fscale or 0) / 1024.0) if fscale else None
        wr = mean(lst, "WRITE_SIZE", (wscale or 0) / 1024.0) if wscale else None
        ai, tc = mean(lst, "SQ_ACTIVE_INST_VALU"), mean(lst, "SQ_THREAD_CYCLES_VALU")
        util = 100.0 * tc / (64.0 * ai) if ai and tc is not None else None
        lb, la = mean(lst, "SQ_LDS_BANK_CONFLICT"), mean(lst, "SQ_LDS_IDX_ACTIVE")
        gui = mean(lst, "GRBM_GUI_ACTIVE")
        raw = 100.0 * la / (256.0 * gui / 8.0) if la is not None and gui else None
        lr = 100.0 * lb / la if la and lb is not None else None
        out.append("| `%s` | %d | %s | %s | %s | %s | %s | %s | %s | %s |" % (
            k, n_launch, fmt(fe, "%.2f"), fmt(wr, "%.2f"), fmt(mean(lst, "SQ_INSTS_VALU"), "%.3g"),
            fmt(mean(lst, "VALUBusy"), "%.1f"), fmt(raw, "%.1f"), fmt(util, "%.1f"), fmt(mean(lst, "LDSBankConflict"), "%.1f"),
            fmt(lr, "%.1f")))
    out.append("")
    traffic = None
    if fscale and wscale and "ext_batched_timed" in labels:
        seg_i = labels.index("ext_batched_timed")
        warm = segs[seg_i].get("warmup", 0)

        def per_batch(key, scale):
            v = []
            for pieces in split:
                for q in _launch_sequences(pieces[1 + seg_i])[warm:]:
                    if any(key in r[4] for r in q):
                        v.append(sum(r[4].get(key, 0.0) for r in q) * scale * 1024.0)
            return v

        fb, wb = per_batch("FETCH_SIZE", fscale), per_batch("WRITE_SIZE", wscale)
        if fb and wb:
            traffic = {"jobs": meta.get("jobs"), "bytes_per_batch": float(np.mean(fb) + np.mean(wb)),
                       "fetch_bytes": float(np.mean(fb)), "write_bytes": float(np.mean(wb)),
                       "fetch_scale": fscale, "write_scale": wscale, "launch_sequences": [len(fb), len(wb)]}
            out.append("Extension batch traffic (every kernel of the batch's launch sequence, mean of the timed "
                       "runs): %.1f MiB fetched + %.1f MiB written = %.1f MiB per batch of %s jobs." %
                       (np.mean(fb) / 2**20, np.mean(wb) / 2**20, (np.mean(fb) + np.mean(wb)) / 2**20,
                        meta.get("jobs")))
            out.append("")
    print("\n".join(out))
    if traffic_out and traffic:
        with open(traffic_out, "w") as f:
            json.dump(traffic, f, indent=1)


def pct(v, f):
    return float(np.percentile(np.asarray(v), f)) if len(v) else None


def call_shapes(ctx, threads, ext_calls, resc_calls, reps, check):
    """The reference's call shapes on one context: `threads` host threads each making extension
    calls of 2000 jobs, then rescue calls of 100 jobs, directly (bwg_extend_batch /
    bwg_ksw_align2_batch on the context's execution lanes) and through the server (submit + wait
    per call).  Returns jobs/s and call-latency percentiles; every result is compared with the same
    call made alone beforehand."""
    ep = bwg.ext_params(5, bwg.default_mat(), *EXT_P)
    kp = bwg.ksw_params(5, bwg.default_mat(), *RESC_P)
    # the calls one at a time first: the reference results, and the mean host phases of a call
    out = {}
    alone_e, alone_r = [], []
    for kind, calls, alone in (("ext", ext_calls, alone_e), ("resc", resc_calls, alone_r)):
        ph = []
        t0 = time.perf_counter()
        for b in calls:
            alone.append(ctx.extend_arrays(b, ep) if kind == "ext" else ctx.align2_arrays(b, kp))
            ph.append(ctx.phases())
        wall = time.perf_counter() - t0
        out["%s_one_thread" % kind] = {
            "calls": len(calls), "jobs_per_call": len(calls[0]), "jobs_per_s": sum(len(b) for b in calls) / wall,
            "phases_ms": {k: round(float(np.mean([p[k] for p in ph])) * 1e3, 3) for k in ph[0]}}
        log("call shapes %s one thread: %.3g jobs/s, mean phases (ms) %s"
            % (kind, out["%s_one_thread" % kind]["jobs_per_s"], out["%s_one_thread" % kind]["phases_ms"]))
    if check:
        ob = oracle()
        mat = bwg.default_mat()
        for b, a in zip(ext_calls[:16], alone_e[:16]):
            assert np.array_equal(a, ob.extend_batch(b, mat, *EXT_P)), "call-shape extension vs oracle"
        for b, a in zip(resc_calls[:32], alone_r[:32]):
            assert np.array_equal(a, ob.align2_batch(b, mat, *RESC_P)), "call-shape rescue vs oracle"
    srv = bwg.Server(ctx, hbm_budget_bytes=64 << 30, slots=3)
    try:
        for kind, calls, alone, params in (("ext", ext_calls, alone_e, ep), ("resc", resc_calls, alone_r, kp)):
            for route in ("lanes", "direct", "server"):
                # lanes: synchronous calls each on an execution lane (coalescing off); direct: the
                # same calls with coalescing on (the default); server: explicit submit + wait
                ctx.set_option(bwg.OPT_COALESCE, 0 if route == "lanes" else 1)
                lat, bad = [], []
                lock = threading.Lock()

                def worker(t):
                    mine = []
                    for r in range(reps):
                        k = (t * reps + r) % len(calls)
                        t0 = time.perf_counter()
                        if route != "server":
                            got = (ctx.extend_arrays(calls[k], params) if kind == "ext"
                                   else ctx.align2_arrays(calls[k], params))
                        else:
                            got = srv.wait(srv.submit(calls[k], params))
                        mine.append(time.perf_counter() - t0)
                        if not np.array_equal(got, alone[k]):
                            bad.append(k)
                    with lock:
                        lat.extend(mine)

                th = [threading.Thread(target=worker, args=(t,)) for t in range(threads)]
                t0 = time.perf_counter()
                for x in th:
                    x.start()
                for x in th:
                    x.join()
                wall = time.perf_counter() - t0
                if bad:
                    log("call shapes %s/%s: %d results differ from the call alone" % (kind, route, len(bad)))
                    sys.exit(1)
                njobs = sum(len(calls[(t * reps + r) % len(calls)]) for t in range(threads) for r in range(reps))
                out["%s_%s" % (kind, route)] = {
                    "threads": threads, "calls": threads * reps, "jobs_per_call": len(calls[0]),
                    "jobs_per_s": njobs / wall, "p50_call_ms": pct(lat, 50) * 1e3, "p99_call_ms": pct(lat, 99) * 1e3}
                log("call shapes %s %s: %d threads x %d calls of %d jobs: %.3g jobs/s, call p50 %.2f ms p99 %.2f ms"
                    % (kind, route, threads, reps, len(calls[0]), njobs / wall, pct(lat, 50) * 1e3,
                       pct(lat, 99) * 1e3))
    finally:
        srv.close()
    return out


def main():
    hi = host_info()
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--jobs", type=int, default=1_000_000)
    ap.add_argument("--rescue-jobs", type=int, default=200_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=hi["threads"])
    ap.add_argument("--forms", default="batched,plain", help="comma list of forms to time")
    ap.add_argument("--only", default="ext,bins,small,resc,side,matesw,global,e2e,calls",
                    help="comma list of workloads to time")
    ap.add_argument("--global-jobs", type=int, default=200_000,
                    help="region-to-CIGAR jobs of the banded global alignment workload")
    ap.add_argument("--matesw-anchors", type=int, default=200_000,
                    help="anchors of the mate-rescue round (4 orientation jobs each)")
    ap.add_argument("--serve-rates", default="1e3,5e4,2e5,2e6,1e7,3e7,6e7,1e8",
                    help="offered loads (jobs/s) of the serving sweep; empty to skip")
    ap.add_argument("--serve-seconds", type=float, default=3.0,
                    help="seconds per offered load (longer at low loads, for at least ~40 chunks)")
    ap.add_argument("--serve-min-chunks", type=int, default=500,
                    help="chunks at the config-4 point (200k jobs/s) that its p50/p99 are taken over")
    ap.add_argument("--call-reps", type=int, default=8, help="calls per thread of the call-shape leg")
    ap.add_argument("--small-jobs", type=int, default=100_000,
                    help="jobs of the small plain-extension workload (0 to skip)")
    ap.add_argument("--serial-bins", action="store_true",
                    help="run a batch's bins one after another on one stream (for counter runs)")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="also run one device-to-device copy of this size (HBM counter calibration)")
    ap.add_argument("--segments-out", default="",
                    help="write the trace segment labels and the per-bin cell tables here (JSON)")
    ap.add_argument("--summarize-trace", nargs=2, metavar=("KERNEL_TRACE_CSV", "SEGMENTS_JSON"),
                    help="summarise a rocprofv3 kernel trace of a bench run (no GPU needed) and exit")
    ap.add_argument("--summarize-counters", nargs="+", metavar="CSV",
                    help="summarise rocprofv3 --pmc runs of one bench command (counter_collection CSVs, then "
                         "its --segments-out JSON) and exit")
    ap.add_argument("--traffic-out", default="",
                    help="with --summarize-counters: write the extension batch's HBM bytes here as JSON")
    args = ap.parse_args()
    if args.summarize_trace:
        summarize_trace(*args.summarize_trace)
        return
    if args.summarize_counters:
        if len(args.summarize_counters) < 2:
            raise SystemExit("--summarize-counters takes counter CSVs and then a segments JSON")
        summarize_counters(args.summarize_counters[:-1], args.summarize_counters[-1], args.traffic_out)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this path has no multi-GPU step (jobs shard across GPUs "
              "with no exchange); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    mat = bwg.default_mat()
    ext_p = bwg.ext_params(5, mat, *EXT_P)
    ksw_p = bwg.ksw_params(5, mat, *RESC_P)
    only = args.only.split(",")
    log("host: %s, %d CPUs online, %d available, CPU legs on %d threads"
        % (hi["cpu_model"], hi["cpus_online"], hi["cpus_available"], args.cpu_threads))
    log("generating %d long-tail extension jobs, %d small extension jobs and %d rescue jobs"
        % (args.jobs, args.small_jobs, args.rescue_jobs))
    ext = bwg.generate(bwg.WORKLOAD_EXT_TAIL, args.jobs, seed=42)
    resc = bwg.generate(bwg.WORKLOAD_RESCUE, args.rescue_jobs, seed=42)
    small = bwg.generate(bwg.WORKLOAD_EXT_SMALL, args.small_jobs, seed=42) if args.small_jobs else None

    # the oracle's per-job cells and ungapped-diagonal hits: the GCUPS and roofline denominators
    cpu = cpu_resc = None
    job_cells = hit = None
    ob = oracle()  # the checker, the cell counts (the GCUPS unit) and the CPU baseline
    if not args.no_cpu or "bins" in only:
        log("oracle over the extension batch on %d threads (cells and the CPU baseline)" % args.cpu_threads)
        t0 = time.perf_counter()
        cpu_out, hit = ob.ungapped_batch(ext, mat, 6, 1, 6, 1, 100)
        idx = np.nonzero(~hit)[0]
        rest = bwg.JobBatch(ext.arena, ext.q_off[idx], ext.q_len[idx], ext.t_off[idx], ext.t_len[idx], ext.h0[idx])
        cpu_out[idx] = ob.extend_batch(rest, mat, *EXT_P, nthreads=args.cpu_threads)
        cpu_dt = time.perf_counter() - t0
        full, job_cells = ob.extend_batch_job_cells(ext, mat, *EXT_P, nthreads=args.cpu_threads)
        assert np.array_equal(full, cpu_out)
        # the SIMD CPU path: the reference's inter-sequence form (16 jobs in an AVX2 register)
        t0 = time.perf_counter()
        simd_out, _ = ob.ungapped_batch(ext, mat, 6, 1, 6, 1, 100)
        simd_out[idx] = ob.simd_extend_batch(rest, mat, *EXT_P, nthreads=args.cpu_threads)
        simd_dt = time.perf_counter() - t0
        assert np.array_equal(simd_out, cpu_out), "SIMD CPU baseline differs from the scalar oracle"
        cpu = {"value": args.jobs / simd_dt, "unit": "jobs/s", "cores": args.cpu_threads, "kind": "port",
               "gcups": float(job_cells.sum()) / simd_dt / 1e9, "cpu_model": hi["cpu_model"],
               "cpus_online": hi["cpus_online"], "cpus_available": hi["cpus_available"],
               "sample": "all %d long-tail extension jobs: the ungapped-diagonal fast path, then the "
                         "reference's inter-sequence SIMD form (AVX2, 16 jobs per register, union band, "
                         "per-lane row epilogue) for the declined jobs, one thread per core, static "
                         "partition" % args.jobs,
               "scalar": {"value": args.jobs / cpu_dt, "gcups": float(job_cells.sum()) / cpu_dt / 1e9,
                          "sample": "the same jobs through the scalar C ksw_extend2"}}
        log("cpu baseline: %.3g jobs/s SIMD, %.3g jobs/s scalar, on %d threads"
            % (cpu["value"], cpu["scalar"]["value"], args.cpu_threads))

    ctx = bwg.Context(0)
    if args.serial_bins:
        ctx.set_option(bwg.OPT_CONCURRENT_BINS, 0)
    copy_gbs = None
    if args.calib_copy_mb:
        copy_gbs = ctx.hbm_copy(args.calib_copy_mb << 20, reps=1)
        SEGMENTS.append(dict(label="setup"))  # the calibration copy's dispatches look like a marker's
        log("calibration copy: hbm_copy_kernel, %d MiB read + %d MiB written per launch, %.0f GB/s"
            % (args.calib_copy_mb, args.calib_copy_mb, copy_gbs))
    res, outs = {}, {}
    ext_bins = resc_bins = None
    resc_exec = None
    forms = [f for f in ((bwg.FORM_BATCHED, "batched"), (bwg.FORM_PLAIN, "plain")) if f[1] in args.forms.split(",")]
    for form, fname in forms:
        ctx.set_form(form)
        if "ext" in only:
            log("extension %s: staging" % fname)
            st = ctx.stage(ext, ext_p)
            st.run()  # first run builds the form's plan and buffers (and tunes)
            mark(ctx, "ext_%s_timed" % fname, warmup=args.warmup)
            per, kern = time_steps(ctx, st, args.steps if form == bwg.FORM_BATCHED else max(1, args.steps // 3),
                                   args.warmup)
            mark(ctx, "ext_%s_after" % fname)
            outs["ext_" + fname] = st.results()
            hits = st.ungapped_hits()
            res["ext_" + fname] = dict(step_s=per, kernel_s=kern, ungapped_hits=hits)
            log("extension %s: %.3f ms/step (kernels %.3f ms), %d jobs resolved by the ungapped fast path"
                % (fname, per * 1e3, kern * 1e3, hits))
            if form == bwg.FORM_BATCHED and "bins" in only and job_cells is not None:
                mark(ctx, "ext_serial_bins")
                rep, tot = serial_bin_times(ctx, st)
                mark(ctx, "ext_serial_after")
                ext_bins = ext_bin_table(rep, st.job_bins(), job_cells, ~hit)
                res["ext_serial"] = dict(kernel_s=tot)
                for row in ext_bins:
                    log("  serial bin %-22s jobs %7d  %8.1f us  frac %s" % (row["kernel"], row["jobs"], row["us"],
                                                                            row.get("frac")))
            st.close()
            if form == bwg.FORM_BATCHED:
                # the same batch with every job through the banded DP, for comparison
                ctx.set_option(bwg.OPT_UNGAPPED, 0)
                st = ctx.stage(ext, ext_p)
                st.run()
                per0, kern0 = time_steps(ctx, st, max(1, args.steps // 2), 1)
                assert np.array_equal(st.results(), outs["ext_batched"]), "fast path changes results"
                st.close()
                ctx.set_option(bwg.OPT_UNGAPPED, 1)
                res["ext_batched_dp_only"] = dict(step_s=per0, kernel_s=kern0)
                log("extension batched, fast path off: %.3f ms/step (kernels %.3f ms)" % (per0 * 1e3, kern0 * 1e3))
        if "small" in only and small is not None:
            st = ctx.stage(small, ext_p)
            st.run()
            per, kern = time_steps(ctx, st, args.steps, args.warmup)
            outs["small_" + fname] = st.results()
            st.close()
            res["small_" + fname] = dict(step_s=per, kernel_s=kern)
            log("small extension %s: %.3f ms/step (kernels %.3f ms)" % (fname, per * 1e3, kern * 1e3))
        if "resc" in only:
            log("rescue %s: staging" % fname)
            st = ctx.stage(resc, ksw_p)
            st.run()
            mark(ctx, "resc_%s_timed" % fname, warmup=1)
            per, kern = time_steps(ctx, st, max(1, args.steps // 2) if form == bwg.FORM_BATCHED else 2, 1)
            mark(ctx, "resc_%s_after" % fname)
            outs["resc_" + fname] = st.results()
            res["resc_" + fname] = dict(step_s=per, kernel_s=kern)
            log("rescue %s: %.3f ms/step (kernels %.3f ms)" % (fname, per * 1e3, kern * 1e3))
            if form == bwg.FORM_BATCHED and "bins" in only:
                mark(ctx, "resc_serial_bins")
                rep, tot = serial_bin_times(ctx, st)
                mark(ctx, "resc_serial_after")
                resc_bins, fwd_pad, rev_cells = resc_bin_table(rep, st.job_bins(), resc, outs["resc_batched"])
                resc_exec = dict(forward_padded_cells=fwd_pad, reverse_cells=rev_cells, serial_kernel_s=tot)
                for row in resc_bins:
                    log("  serial launch %-28s jobs %7d  %8.1f us  frac %s" % (row["kernel"], row["jobs"], row["us"],
                                                                               row.get("frac")))
            st.close()

    # rescue: both forms agree, and the oracle (config 2's minsc 20) agrees with them
    if "resc_batched" in outs and "resc_plain" in outs and not np.array_equal(outs["resc_batched"], outs["resc_plain"]):
        log("rescue forms disagree")
        sys.exit(1)
    resc_head = "resc_batched" if "resc_batched" in outs else ("resc_plain" if "resc_plain" in outs else None)
    if resc_head and not (args.no_cpu and args.no_check):
        log("oracle over the %d rescue jobs on %d threads" % (args.rescue_jobs, args.cpu_threads))
        t0 = time.perf_counter()
        want = ob.align2_batch(resc, mat, *RESC_P, nthreads=args.cpu_threads)
        dt = time.perf_counter() - t0
        ok = np.array_equal(want, outs[resc_head])
        log("rescue bit-exact vs oracle (minsc %d): %s" % (RESC_P[4], ok))
        if not ok:
            bad = np.nonzero(want != outs[resc_head])[0]
            log("rescue mismatches: %d, first %d" % (bad.size, bad[0]))
            sys.exit(1)
        nominal = float(np.sum(resc.q_len.astype(np.int64) * resc.t_len))
        t0 = time.perf_counter()
        simd = ob.simd_align2_batch(resc, mat, *RESC_P, nthreads=args.cpu_threads)
        sdt = time.perf_counter() - t0
        if not np.array_equal(simd, want):
            log("SIMD CPU rescue baseline differs from the scalar oracle")
            sys.exit(1)
        cpu_resc = {"value": args.rescue_jobs / sdt, "unit": "jobs/s", "cores": args.cpu_threads, "kind": "port",
                    "gcell_s": nominal / sdt / 1e9,
                    "sample": "all %d rescue jobs: the reference's inter-sequence SIMD form of batched "
                              "ksw_align2 (AVX2, 16 jobs per register; forward pass, tracker, reverse pass), "
                              "one thread per core" % args.rescue_jobs,
                    "scalar": {"value": args.rescue_jobs / dt, "gcell_s": nominal / dt / 1e9,
                               "sample": "the same jobs through the scalar C batched ksw_align2"}}
        log("cpu rescue baseline: %.3g jobs/s SIMD, %.3g jobs/s scalar" % (cpu_resc["value"], args.rescue_jobs / dt))

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
            per, kern = time_steps(ctx, st, max(1, args.steps // 2), 1)
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
        per, kern = time_steps(ctx, st, max(1, args.steps // 2), 1)
        win, out = st.windows(), st.results()
        st.close()
        valid = (win["flags"] & bwg.WIN_VALID) != 0
        cells = float(np.sum((win["re"] - win["rb"])[valid].astype(np.int64) *
                             rnd.anchors["l_ms"].repeat(4)[valid].astype(np.int64)))
        matesw = dict(step_s=per, kernel_s=kern, jobs=int(valid.sum()), cells=cells)
        log("mate rescue: %.3f ms/step (kernels %.3f ms), %d windows aligned, %.0f G cells/s"
            % (per * 1e3, kern * 1e3, matesw["jobs"], cells / per / 1e9))
        if not args.no_check:
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
        per, kern = time_steps(ctx, st, max(1, args.steps // 2), 1)
        gscore = st.results()
        gn, gc = st.cigars()
        st.close()
        qg, tg = gb.q_len.astype(np.int64), gb.t_len.astype(np.int64)
        gcells = float(np.sum(np.minimum(qg, 2 * gw.astype(np.int64) + 1) * tg))  # upper bound: full band rows
        glb = dict(step_s=per, kernel_s=kern, cells=gcells)
        log("global alignment: %.3f ms/step (kernels %.3f ms), %.3g jobs/s" % (per * 1e3, kern * 1e3,
                                                                                 args.global_jobs / per))
        if not args.no_check:
            ws, wn, wc = ob.global_batch(gb, gw, mat, 6, 1, 6, 1, nthreads=args.cpu_threads)
            ok = np.array_equal(ws, gscore) and np.array_equal(wn.astype(np.uint32), gn) and np.array_equal(wc, gc)
            log("global alignment: bit-exact vs oracle: %s" % ok)
            if not ok:
                sys.exit(1)
    e2e = e2e_phases = None
    if "e2e" in only and "ext" in only:
        # PCIe-inclusive: the synchronous entry point from host buffers (stage, plan, run, copy
        # back), with the host time of each phase of the last call
        ctx.set_form(bwg.FORM_BATCHED)
        ctx.extend_arrays(ext, ext_p)
        t0 = time.perf_counter()
        reps = 3
        for _ in range(reps):
            got = ctx.extend_arrays(ext, ext_p)
        e2e = args.jobs * reps / (time.perf_counter() - t0)
        e2e_phases = {k: round(v * 1e3, 3) for k, v in ctx.phases().items()}
        e2e_phases["device_kernels"] = round(ctx.stat(bwg.STAT_LAST_KERNEL_NS) * 1e-6, 3)
        if "ext_batched" in outs and not np.array_equal(got, outs["ext_batched"]):
            log("synchronous call differs from the staged batch")
            sys.exit(1)
        log("extension from host buffers (PCIe-inclusive): %.3g jobs/s; phases of the last call (ms): %s"
            % (e2e, e2e_phases))
    calls = None
    if "calls" in only:
        ctx.set_form(bwg.FORM_BATCHED)
        threads = 16
        ext_calls = [bwg.generate(bwg.WORKLOAD_EXT_TAIL, 2000, seed=1000 + k) for k in range(64)]
        resc_calls = [bwg.generate(bwg.WORKLOAD_RESCUE, 100, seed=2000 + k) for k in range(128)]
        calls = call_shapes(ctx, threads, ext_calls, resc_calls, args.call_reps, not args.no_check)
    serving = []
    rates = [float(x) for x in args.serve_rates.split(",") if x.strip()]
    if rates:
        ctx.set_form(bwg.FORM_BATCHED)
        srv = bwg.Server(ctx, hbm_budget_bytes=64 << 30, slots=3)
        for rate in rates:
            # chunks average ~15k jobs: the config-4 point (200k jobs/s) runs for at least
            # serve_min_chunks chunks; the others for ~40 chunks within 20 s, and at least
            # serve_seconds (so the lowest rates complete only a few chunks: their count is reported)
            if abs(rate - 2e5) < 1:
                secs = args.serve_min_chunks * 15500.0 / rate
            else:
                secs = max(args.serve_seconds, min(20.0, 40 * 15000.0 / rate))
            r = bwg.serve_run(ctx, srv, rate, producers=8, duration_s=secs, seed=42,
                              use_ref=True, pool_chunks=64, verify=True)
            r["offered_jobs_per_s"] = rate
            r["seconds"] = secs
            serving.append(r)
            log("serving at %.3g jobs/s for %.0f s: p50 %.2f ms p99 %.2f ms over %d chunks, completed %.3g jobs/s, "
                "%d mismatched" % (rate, secs, r["p50_ms"], r["p99_ms"], r["chunks"], r["jobs_per_s"],
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
    if cpu and not args.no_check and head in outs:
        ok = np.array_equal(cpu_out, outs[head])
        log("extension bit-exact vs oracle: %s" % ok)
        if not ok:
            bad = np.nonzero(cpu_out != outs[head])[0]
            log("mismatches: %d, first %d" % (bad.size, bad[0]))
            sys.exit(1)
    small_cells = None
    small_key = "small_batched" if "small_batched" in outs else "small_plain"
    if ob is not None and small_key in outs:
        s_full, s_cells = ob.extend_batch_job_cells(small, mat, *EXT_P, nthreads=args.cpu_threads)
        small_cells = int(s_cells.sum())
        ok = np.array_equal(s_full, outs[small_key])
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
    if job_cells is not None:
        cells = int(job_cells.sum())
        dp_cells = int(job_cells[~hit].sum())
        line["gcups"] = cells / h["step_s"] / 1e9
        line["gcups_kernel"] = cells / h["kernel_s"] / 1e9
        line["cells"] = cells
        # the DP kernels evaluate the cells of the jobs the ungapped fast path declines
        if h.get("ungapped_hits"):
            assert h["ungapped_hits"] == int(hit.sum()), "device and oracle fast paths disagree"
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
        if "ext_batched_timed" in [s["label"] for s in SEGMENTS] and head == "ext_batched":
            segment("ext_batched_timed")["roofline"] = dict(
                ops=(dp_cells if h.get("ungapped_hits") else cells) * K_EXT, peak=VALU_PEAK_TOPS,
                what="int32 VALU; 16 ops per DP cell of the jobs the ungapped pre-pass declines",
                bench_frac=ops / VALU_PEAK_TOPS)
        if "ext_batched_dp_only" in res:
            line["batched_dp_only_roofline_frac"] = (cells * K_EXT / res["ext_batched_dp_only"]["kernel_s"] / 1e12
                                                     / VALU_PEAK_TOPS)
    if ext_bins:
        line["ext_bins_serial"] = ext_bins
        line["ext_serial_kernel_ms"] = res["ext_serial"]["kernel_s"] * 1e3
    if "ext_batched_dp_only" in res:
        line["batched_dp_only_jobs_per_s"] = args.jobs / res["ext_batched_dp_only"]["step_s"]
    if "ext_plain" in res:
        line["plain_jobs_per_s"] = args.jobs / res["ext_plain"]["step_s"]
        # batched form with the ungapped pre-pass (value) and without it, each over the plain form
        line["speedup_vs_plain"] = jobs_s / line["plain_jobs_per_s"]
        if "batched_dp_only_jobs_per_s" in line:
            line["speedup_vs_plain_dp_only"] = line["batched_dp_only_jobs_per_s"] / line["plain_jobs_per_s"]
    if e2e:
        line["e2e_jobs_per_s"] = e2e
        line["e2e_phases_ms"] = e2e_phases
    if calls:
        line["call_shapes"] = calls
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
            if small_cells:
                line["small_%s_gcups" % f] = small_cells / r["step_s"] / 1e9
    nominal = float(np.sum(resc.q_len.astype(np.int64) * resc.t_len))
    for f in ("batched", "plain"):
        if "resc_" + f in res:
            r = res["resc_" + f]
            line["rescue_%s_jobs_per_s" % f] = args.rescue_jobs / r["step_s"]
            line["rescue_%s_gcell_s" % f] = nominal / r["step_s"] / 1e9
    if "resc_batched" in res:
        r = res["resc_batched"]
        # forward-pass nominal cells x 13 over the whole rescue launch sequence (forward, tracker,
        # reverse); every forward bin of this workload is the packed 16-bit kernel
        packed = resc_bins is not None and all(b["kernel"].startswith("ksw_pk") for b in resc_bins if "cells" in b
                                               and "reverse" not in b["kernel"])
        peak = VALU_PK16_PEAK_TOPS if packed else VALU_PEAK_TOPS
        ach = nominal * K_RESC / r["kernel_s"] / 1e12
        line["rescue_roofline"] = {"bound": "valu", "achieved": ach, "peak": peak, "unit": "Tops/s",
                                   "frac": ach / peak, "cells": nominal,
                                   "arith": "packed 16-bit" if packed else "int32"}
        if resc_exec:
            line["rescue_roofline"].update(resc_exec)
            line["rescue_roofline"]["executed_cells"] = resc_exec["forward_padded_cells"] + resc_exec["reverse_cells"]
        if resc_bins:
            line["rescue_bins_serial"] = resc_bins
        if "resc_batched_timed" in [s["label"] for s in SEGMENTS]:
            segment("resc_batched_timed")["roofline"] = dict(
                ops=nominal * K_RESC, peak=peak, bench_frac=ach / peak,
                what="%s VALU; 13 ops per nominal qlen x tlen cell" % line["rescue_roofline"]["arith"])
    if serving:
        at_cfg = min(serving, key=lambda r: abs(r["offered_jobs_per_s"] - 2e5))
        line["p50_chunk_latency_ms"] = at_cfg["p50_ms"]
        line["p99_chunk_latency_ms"] = at_cfg["p99_ms"]
        line["latency_chunks"] = at_cfg["chunks"]
        line["serving"] = [{k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()}
                           for r in serving]
    rep = ctx.tune_report()
    if rep:
        line["tuning"] = rep.splitlines()
    if args.segments_out:
        with open(args.segments_out, "w") as f:
            json.dump({"segments": SEGMENTS, "ext_bins": ext_bins, "rescue_bins": resc_bins,
                       "steps": args.steps, "jobs": args.jobs, "rescue_jobs": args.rescue_jobs,
                       "calib_copy_mb": args.calib_copy_mb}, f, indent=1)
    if cpu:
        line["cpu_baseline"] = cpu
    if cpu_resc:
        line["cpu_baseline_rescue"] = cpu_resc
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
