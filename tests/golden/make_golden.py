"""Write the golden fixtures in this directory: job inputs and the oracle's expected outputs.

    python3 tests/golden/make_golden.py

Fixtures (numpy .npz, no pickled objects):

* ``ext_workloads.npz``: 400 jobs each of the small plain-extension workload (query 1-130 from
  150 bp reads, target qlen+100, h0 19-150) and of the long-tail batched workload (reads
  100/150/250, query 70% <30 / 25% 30-150 / 5% 150-250, h0 19-250), both seed 42, scored
  A=1 B=4 O=6 E=1 w=100 zdrop=100 end_bonus=5.
* ``ext_scorings.npz``: jobs of the reference's acceptance-gate sweeps (LCG-generated, varying
  band, z-drop, end bonus and gap penalties per batch), flattened with per-job scoring.
* ``rescue.npz``: 100 jobs of the mate-rescue workload (150 bp reads in 600-1200 bp windows, 30%
  random, minsc 20) and 300 jobs of the rescue gate sweep under its two scorings.

Expected outputs come from the C oracle (oracle/ksw_oracle.c), which the oracle tests pin against
the reference's known answers, brute force and closed forms; the fixtures guard the oracle and the
kernels against regressions and are regenerated only when an input generator changes.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "bwa-mem4_amd"))

import bwg  # noqa: E402
import lcg_gates as G  # noqa: E402
import oracle_binding as ob  # noqa: E402

EXT_FIELDS = ("score", "qle", "tle", "gtle", "gscore", "max_off")
KSW_FIELDS = ("score", "qb", "qe", "tb", "te", "score2", "te2")


def _as2d(rec, fields):
    return np.stack([rec[f] for f in fields], axis=1).astype(np.int32)


def _pack(batch, prefix, d):
    for name in ("arena", "q_off", "q_len", "t_off", "t_len", "h0"):
        d[prefix + name] = getattr(batch, name)


def main():
    mat = bwg.default_mat()
    d = {}
    for name, kind in (("small_", bwg.WORKLOAD_EXT_SMALL), ("tail_", bwg.WORKLOAD_EXT_TAIL)):
        b = bwg.generate(kind, 400, seed=42)
        _pack(b, name, d)
        d[name + "want"] = _as2d(ob.extend_batch(b, mat, 6, 1, 6, 1, 100, 5, 100), EXT_FIELDS)
    np.savez_compressed(os.path.join(HERE, "ext_workloads.npz"), **d)

    # gate sweeps: every 4th batch of the two randomized sweeps, flattened with per-job scoring
    qs, ts, h0s, sc, want = [], [], [], [], []
    for gen in (G.ext_batch_rounds, G.ext_read_sized):
        for r, case in enumerate(gen()):
            if r % 4:
                continue
            p = (case["o_del"], case["e_del"], case["o_ins"], case["e_ins"], case["w"],
                 case["end_bonus"], case["zdrop"])
            for q, t, h0 in zip(case["queries"], case["targets"], case["h0s"]):
                qs.append(q)
                ts.append(t)
                h0s.append(h0)
                sc.append(p)
                want.append(ob.extend2(q, t, case["mat"], *p, h0))
    b = bwg.JobBatch.from_pairs(qs, ts, h0s)
    d = {}
    _pack(b, "", d)
    d["scoring"] = np.asarray(sc, dtype=np.int32)  # o_del e_del o_ins e_ins w end_bonus zdrop
    d["want"] = np.asarray([tuple(w) for w in want], dtype=np.int32)
    np.savez_compressed(os.path.join(HERE, "ext_scorings.npz"), **d)

    d = {}
    b = bwg.generate(bwg.WORKLOAD_RESCUE, 100, seed=42)
    _pack(b, "work_", d)
    d["work_want"] = _as2d(ob.align2_batch(b, mat, 6, 1, 6, 1, 20, 1), KSW_FIELDS)
    gq, gt = G.rescue_jobs()
    b = bwg.JobBatch.from_pairs(gq[:300], gt[:300])
    _pack(b, "gate_", d)
    for a, bb, minsc in G.RESCUE_SCORINGS:
        m = G.scmat(a, bb)
        d["gate_want_%d_%d_%d" % (a, bb, minsc)] = _as2d(
            ob.align2_batch(b, m, 6, 1, 6, 1, minsc, a), KSW_FIELDS)
    np.savez_compressed(os.path.join(HERE, "rescue.npz"), **d)
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
