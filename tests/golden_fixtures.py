"""Loaders for the golden fixtures in tests/golden (inputs + expected outputs, see make_golden.py)."""
import os

import numpy as np

import bwg

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
EXT_FIELDS = ("score", "qle", "tle", "gtle", "gscore", "max_off")
KSW_FIELDS = ("score", "qb", "qe", "tb", "te", "score2", "te2")


def load(name):
    with np.load(os.path.join(HERE, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def batch(d, prefix=""):
    return bwg.JobBatch(d[prefix + "arena"], d[prefix + "q_off"], d[prefix + "q_len"],
                        d[prefix + "t_off"], d[prefix + "t_len"], d[prefix + "h0"])


def as2d(rec, fields):
    return np.stack([rec[f] for f in fields], axis=1).astype(np.int32)


def scoring_groups(d):
    """ext_scorings.npz: (scoring tuple, job indices) per distinct scoring."""
    sc = d["scoring"]
    keys = {}
    for k, row in enumerate(map(tuple, sc.tolist())):
        keys.setdefault(row, []).append(k)
    return [(k, np.asarray(v)) for k, v in keys.items()]
