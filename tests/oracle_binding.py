"""ctypes binding to the CPU oracle (oracle/liboracle_ksw.so).

Test infrastructure: only the tests, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg
load this, as the checker.  The product library never does.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# BWG_ORACLE_SO selects another build of the same sources (the sanitizer build, oracle/Makefile)
ORACLE_SO = os.environ.get("BWG_ORACLE_SO") or os.path.join(ROOT, "oracle", "liboracle_ksw.so")


class ExtResult(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("score", "qle", "tle", "gtle", "gscore", "max_off")]

    def tup(self):
        return (self.score, self.qle, self.tle, self.gtle, self.gscore, self.max_off)


class KswResult(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("score", "qb", "qe", "tb", "te", "score2", "te2")]

    def tup(self):
        return (self.score, self.qb, self.qe, self.tb, self.te, self.score2, self.te2)


class FwdResult(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("score", "te", "qe", "score2", "te2")]

    def tup(self):
        return (self.score, self.te, self.qe, self.score2, self.te2)


EXT_DTYPE = np.dtype([(n, "<i4") for n in ("score", "qle", "tle", "gtle", "gscore", "max_off")])
KSW_DTYPE = np.dtype([(n, "<i4") for n in ("score", "qb", "qe", "tb", "te", "score2", "te2")])

_P = C.c_void_p
_I = C.c_int32
_lib = None


def build_oracle():
    import subprocess
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build_oracle()
        L = C.CDLL(ORACLE_SO)
        L.orc_ksw_extend2.restype = ExtResult
        L.orc_ksw_extend2.argtypes = [_I, _P, _I, _P, _I, _P, _I, _I, _I, _I, _I, _I, _I, _I, _P]
        L.orc_ksw_local_fwd.restype = FwdResult
        L.orc_ksw_local_fwd.argtypes = [_I, _P, _I, _P, _I, _P, _I, _I, _I, _I, _I, _I, _I, _I, _P]
        L.orc_enum_prefix.restype = None
        L.orc_enum_prefix.argtypes = [_I, _P, _I, _P, _P, _I, _I, _I, _I, _I, _I, _P, _P, _P]
        L.orc_ksw_align2.restype = KswResult
        L.orc_ksw_align2.argtypes = [_I, _P, _I, _P, _I, _P, _I, _I, _I, _I, _I, _I, _I, _P]
        L.orc_batched_align2_job.restype = KswResult
        L.orc_batched_align2_job.argtypes = [_I, _P, _I, _P, _I, _P, _I, _I, _I, _I, _I, _I, _P]
        for f in ("orc_clamp_band", "orc_clamp_band_int"):
            getattr(L, f).restype = _I
            getattr(L, f).argtypes = [_I] * 8
        L.orc_lanes_for.restype = _I
        L.orc_lanes_for.argtypes = [_I, _I]
        L.orc_padded_qlen.restype = _I
        L.orc_padded_qlen.argtypes = [_I, _I]
        L.orc_tracker_push_row.restype = None
        L.orc_tracker_push_row.argtypes = [_P, _I, _I, _I]
        L.orc_tracker_finish.restype = None
        L.orc_tracker_finish.argtypes = [_P, _I, _I, _I, _P, _P]
        L.orc_extend_batch.restype = None
        L.orc_extend_batch.argtypes = [C.c_size_t, _P, _P, _P, _P, _P, _P, _I, _P,
                                       _I, _I, _I, _I, _I, _I, _I, _P, _P, C.c_int]
        L.orc_align2_batch.restype = None
        L.orc_align2_batch.argtypes = [C.c_size_t, _P, _P, _P, _P, _P, _I, _P,
                                       _I, _I, _I, _I, _I, _I, C.c_int, _P, _P, C.c_int]
        L.orc_lockstep_extend.restype = None
        L.orc_lockstep_extend.argtypes = [C.c_size_t, _P, _P, _P, _P, _P, _I, _P,
                                          _I, _I, _I, _I, _I, _I, _I, _I, _P]
        L.orc_lockstep_local_fwd.restype = None
        L.orc_lockstep_local_fwd.argtypes = [C.c_size_t, _P, _P, _P, _P, _P, _P, _I, _P,
                                             _I, _I, _I, _I, _I, _I, _P]
        L.orc_brute_extend_unbanded.restype = _I
        L.orc_brute_extend_unbanded.argtypes = [_I, _P, _I, _P, _I, _P, _I, _I, _I, _I, _I]
        L.orc_brute_gotoh_local.restype = _I
        L.orc_brute_gotoh_local.argtypes = [_I, _P, _I, _P, _I, _P, _I, _I, _I, _I]
        L.orc_brute_enum_local.restype = None
        L.orc_brute_enum_local.argtypes = [_I, _P, _I, _P, _I, _P, _I, _I, _I, _I, _P, _P, _P]
        L.orc_extend_batch_cells.restype = None
        L.orc_extend_batch_cells.argtypes = [C.c_size_t, _P, _P, _P, _P, _P, _P, _I, _P,
                                             _I, _I, _I, _I, _I, _I, _I, _P, _P, C.c_int]
        L.orc_ungapped_batch.restype = None
        L.orc_ungapped_batch.argtypes = [C.c_size_t, _P, _P, _P, _P, _P, _P, _I, _P, _I, _I, _I, _I, _I,
                                         _P, _P]
        L.orc_ungapped_xmax.restype = _I
        L.orc_ungapped_xmax.argtypes = [_P, _I, _I, _I, _I]
        L.orc_extend_side.restype = ExtResult
        L.orc_extend_side.argtypes = [_I, _P, _I, _P, _I, _P, _I, _I, _I, _I, _I, _I, _I, _I, _I, _I, _P]
        L.orc_infer_dir.restype = None
        L.orc_infer_dir.argtypes = [C.c_int64, C.c_int64, C.c_int64, _P, _P]
        L.orc_pos2rid.restype = _I
        L.orc_pos2rid.argtypes = [C.c_int64, _P, _I, C.c_int64]
        L.orc_fetch_window.restype = None
        L.orc_fetch_window.argtypes = [C.c_int64, _P, _P, _I, C.c_int64, C.c_int64, C.c_int64, _P, _P, _P]
        L.orc_ref_base.restype = C.c_uint8
        L.orc_ref_base.argtypes = [_P, C.c_int64, C.c_int64]
        L.orc_matesw_batch.restype = None
        L.orc_matesw_batch.argtypes = [C.c_int64, _P, _P, _P, _I, _P, _P, _P, _I, C.c_size_t, _P, _P, _P, _P,
                                       _I, _I, _I, _I, _I, _I, _P, _P, C.c_int]
        L.orc_ksw_global2.restype = _I
        L.orc_ksw_global2.argtypes = [_I, _P, _I, _P, _I, _P, _I, _I, _I, _I, _I, _P, _P]
        L.orc_global_batch.restype = None
        L.orc_global_batch.argtypes = [C.c_size_t, _P, _P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _P, _P, _P, _P,
                                       C.c_int]
        L.orc_brute_global.restype = _I
        L.orc_brute_global.argtypes = [_I, _P, _I, _P, _P, _I, _I, _I, _I]
        L.orc_simd_extend_batch.restype = None
        L.orc_simd_extend_batch.argtypes = [C.c_size_t, _P, _P, _P, _P, _P, _P, _I, _P,
                                            _I, _I, _I, _I, _I, _I, _I, _P, C.c_int]
        L.orc_simd_align2_batch.restype = None
        L.orc_simd_align2_batch.argtypes = [C.c_size_t, _P, _P, _P, _P, _P, _I, _P,
                                            _I, _I, _I, _I, _I, _I, _P, C.c_int]
        L.orc_run_side.restype = None
        L.orc_run_side.argtypes = [C.c_size_t, _P, _P, _P, _P, _P, _P, _P, _I, _P, _I, _I, _I, _I, _I, _I, _I,
                                   _I, _P, _P, C.c_int]
        _lib = L
    return _lib


def _buf(seq):
    a = np.ascontiguousarray(np.asarray(seq, dtype=np.uint8))
    return a, a.ctypes.data


def _mat(mat):
    a = np.ascontiguousarray(np.asarray(mat, dtype=np.int8))
    assert a.size == 25
    return a, a.ctypes.data


def scmat(a=1, b=4):
    """bwa's 5x5 matrix: match a, mismatch -b, N row/column -1 (crates/bwa-core/src/opt.rs:445)."""
    m = np.full((5, 5), -1, dtype=np.int8)
    for i in range(4):
        for j in range(4):
            m[i, j] = a if i == j else -b
    return m.reshape(-1)


def extend2(q, t, mat, o_del, e_del, o_ins, e_ins, w, end_bonus, zdrop, h0, want_cells=False):
    qa, qp = _buf(q)
    ta, tp = _buf(t)
    ma, mp = _mat(mat)
    cells = C.c_int64(0)
    r = lib().orc_ksw_extend2(len(qa), qp, len(ta), tp, 5, mp, o_del, e_del, o_ins, e_ins, w,
                              end_bonus, zdrop, h0, C.addressof(cells))
    return (r.tup(), cells.value) if want_cells else r.tup()


def extend_side(q, t, mat, o_del, e_del, o_ins, e_ins, w, end_bonus, zdrop, h0, prev0, max_band_try=2):
    """One job's band tries (extend_side): (result tuple, accepted band)."""
    qa, qp = _buf(q)
    ta, tp = _buf(t)
    ma, mp = _mat(mat)
    wo = C.c_int32(0)
    r = lib().orc_extend_side(len(qa), qp, len(ta), tp, 5, mp, o_del, e_del, o_ins, e_ins, w, end_bonus, zdrop,
                              h0, prev0, max_band_try, C.addressof(wo))
    return r.tup(), wo.value


def run_side(batch, prev, mat, o_del, e_del, o_ins, e_ins, w, end_bonus, zdrop, max_band_try=2, nthreads=8):
    """The batched band-retry rounds (run_side): (results, accepted band per job)."""
    n = len(batch.q_len)
    out = np.zeros(n, dtype=EXT_DTYPE)
    w_out = np.zeros(n, dtype=np.int32)
    prev = np.ascontiguousarray(prev, dtype=np.int32)
    ma, mp = _mat(mat)
    lib().orc_run_side(n, batch.arena.ctypes.data, batch.q_off.ctypes.data, batch.q_len.ctypes.data,
                       batch.t_off.ctypes.data, batch.t_len.ctypes.data, batch.h0.ctypes.data, prev.ctypes.data,
                       5, mp, o_del, e_del, o_ins, e_ins, w, end_bonus, zdrop, max_band_try, out.ctypes.data,
                       w_out.ctypes.data, nthreads)
    return out, w_out


WINDOW_DTYPE = np.dtype([("rb", "<i8"), ("re", "<i8"), ("rid", "<i4"), ("flags", "<i4")])


def infer_dir(l_pac, b1, b2):
    d, dist = C.c_int32(0), C.c_int64(0)
    lib().orc_infer_dir(l_pac, b1, b2, C.addressof(d), C.addressof(dist))
    return d.value, dist.value


def _contig_arrays(contigs):
    off = np.ascontiguousarray(contigs["offset"], dtype=np.int64)
    ln = np.ascontiguousarray(contigs["len"], dtype=np.int32)
    return off, ln


def pos2rid(l_pac, contigs, pos_f):
    off, _ = _contig_arrays(contigs)
    return lib().orc_pos2rid(l_pac, off.ctypes.data, len(off), pos_f)


def fetch_window(l_pac, contigs, rb, mid, re):
    off, ln = _contig_arrays(contigs)
    a, b, r = C.c_int64(0), C.c_int64(0), C.c_int32(0)
    lib().orc_fetch_window(l_pac, off.ctypes.data, ln.ctypes.data, len(off), rb, mid, re, C.addressof(a),
                           C.addressof(b), C.addressof(r))
    return a.value, b.value, r.value


def ref_bases(fwd, rb, re):
    fwd = np.ascontiguousarray(fwd, dtype=np.uint8)
    return np.array([lib().orc_ref_base(fwd.ctypes.data, len(fwd), p) for p in range(rb, re)], dtype=np.uint8)


def matesw_batch(rnd, mat, o_del, e_del, o_ins, e_ins, minsc, max_sc, nthreads=8):
    """One mate-rescue round (RescueRound): (windows, results), 4 per anchor."""
    n = len(rnd.anchors)
    win = np.zeros(4 * n, dtype=WINDOW_DTYPE)
    out = np.zeros(4 * n, dtype=KSW_DTYPE)
    off, ln = _contig_arrays(rnd.contigs)
    low = np.ascontiguousarray(rnd.pes["low"], dtype=np.int32)
    high = np.ascontiguousarray(rnd.pes["high"], dtype=np.int32)
    failed = np.ascontiguousarray(rnd.pes["failed"], dtype=np.int32)
    ma, mp = _mat(mat)
    lib().orc_matesw_batch(rnd.l_pac, rnd.fwd.ctypes.data, off.ctypes.data, ln.ctypes.data, len(off),
                           low.ctypes.data, high.ctypes.data, failed.ctypes.data, rnd.min_seed_len, n,
                           rnd.anchors.ctypes.data, rnd.ma_rb.ctypes.data, rnd.arena.ctypes.data, mp, o_del, e_del,
                           o_ins, e_ins, minsc, max_sc, win.ctypes.data, out.ctypes.data, nthreads)
    return win, out


def global2(q, t, mat, o_del, e_del, o_ins, e_ins, w):
    """ksw_global2: (score, CIGAR words len << 4 | op, op 0 = M, 1 = I, 2 = D)."""
    qa, qp = _buf(q)
    ta, tp = _buf(t)
    ma, mp = _mat(mat)
    cig = np.zeros(len(qa) + len(ta) + 2, dtype=np.uint32)
    n = C.c_int32(0)
    sc = lib().orc_ksw_global2(len(qa), qp, len(ta), tp, 5, mp, o_del, e_del, o_ins, e_ins, w, cig.ctypes.data,
                               C.addressof(n))
    return sc, [int(x) for x in cig[:n.value]]


def global_batch(batch, w, mat, o_del, e_del, o_ins, e_ins, nthreads=8):
    """ksw_global2 over a batch with per-job bands: (scores, n_cigar, cigar words concatenated)."""
    n = len(batch.q_len)
    w = np.ascontiguousarray(w, dtype=np.int32)
    cap = batch.q_len.astype(np.uint64) + batch.t_len + 2
    off = np.zeros(n, dtype=np.uint64)
    if n:
        off[1:] = np.cumsum(cap)[:-1]
    buf = np.zeros(int(cap.sum()) + 1, dtype=np.uint32)
    score = np.zeros(n, dtype=np.int32)
    ncig = np.zeros(n, dtype=np.int32)
    ma, mp = _mat(mat)
    lib().orc_global_batch(n, batch.arena.ctypes.data, batch.q_off.ctypes.data, batch.q_len.ctypes.data,
                           batch.t_off.ctypes.data, batch.t_len.ctypes.data, w.ctypes.data, mp, o_del, e_del, o_ins,
                           e_ins, score.ctypes.data, off.ctypes.data, buf.ctypes.data, ncig.ctypes.data, nthreads)
    parts = [buf[int(off[k]):int(off[k]) + int(ncig[k])] for k in range(n)]
    return score, ncig, (np.concatenate(parts) if parts else np.zeros(0, np.uint32))


ENUM_NONE = -(1 << 30)


def enum_prefix(q, t, mat, o_del, e_del, o_ins, e_ins, w, origin):
    """Best score of every alignment prefix from the origin ending at each cell with an aligned
    pair / a deletion / an insertion (orc_enum_prefix: explicit path enumeration, no recurrence).
    Returns three (tlen, qlen) int arrays with ENUM_NONE where no such prefix exists."""
    qa, qp = _buf(q)
    ta, tp = _buf(t)
    ma, mp = _mat(mat)
    out = [np.zeros((len(ta), len(qa)), dtype=np.int32) for _ in range(3)]
    lib().orc_enum_prefix(len(qa), qp, len(ta), tp, mp, o_del, e_del, o_ins, e_ins, w, origin,
                          out[0].ctypes.data, out[1].ctypes.data, out[2].ctypes.data)
    return out


def brute_global(q, t, mat, o_del, e_del, o_ins, e_ins):
    qa, qp = _buf(q)
    ta, tp = _buf(t)
    ma, mp = _mat(mat)
    return lib().orc_brute_global(len(qa), qp, len(ta), tp, mp, o_del, e_del, o_ins, e_ins)


def local_fwd(q, t, mat, o_del, e_del, o_ins, e_ins, minsc, endsc, max_sc, lanes):
    qa, qp = _buf(q)
    ta, tp = _buf(t)
    ma, mp = _mat(mat)
    return lib().orc_ksw_local_fwd(len(qa), qp, len(ta), tp, 5, mp, o_del, e_del, o_ins, e_ins,
                                   minsc, endsc, max_sc, lanes, None).tup()


def align2(q, t, mat, o_del, e_del, o_ins, e_ins, minsc, max_sc, lanes):
    qa, qp = _buf(q)
    ta, tp = _buf(t)
    ma, mp = _mat(mat)
    return lib().orc_ksw_align2(len(qa), qp, len(ta), tp, 5, mp, o_del, e_del, o_ins, e_ins,
                                minsc, max_sc, lanes, None).tup()


def batched_align2_job(q, t, mat, o_del, e_del, o_ins, e_ins, minsc, max_sc):
    qa, qp = _buf(q)
    ta, tp = _buf(t)
    ma, mp = _mat(mat)
    return lib().orc_batched_align2_job(len(qa), qp, len(ta), tp, 5, mp, o_del, e_del, o_ins,
                                        e_ins, minsc, max_sc, None).tup()


class Tracker:
    """The suboptimal tracker, driven row by row (crates/bwa-extend/src/sw.rs:789-852)."""

    def __init__(self, cap=4096):
        self.score = (C.c_int32 * cap)()
        self.row = (C.c_int32 * cap)()

        class T(C.Structure):
            _fields_ = [("score", _P), ("row", _P), ("n", C.c_int32)]

        self.t = T(C.addressof(self.score), C.addressof(self.row), 0)

    def push_row(self, row, row_max, minsc):
        lib().orc_tracker_push_row(C.addressof(self.t), row, row_max, minsc)

    def entries(self):
        return [(self.score[k], self.row[k]) for k in range(self.t.n)]

    def finish(self, gmax, te, max_sc):
        s2, t2 = C.c_int32(0), C.c_int32(0)
        lib().orc_tracker_finish(C.addressof(self.t), gmax, te, max_sc, C.addressof(s2), C.addressof(t2))
        return s2.value, t2.value


def extend_batch(batch, mat, o_del, e_del, o_ins, e_ins, w, end_bonus, zdrop, nthreads=8,
                 want_cells=False):
    """batch: a JobBatch-like object with arena/q_off/q_len/t_off/t_len/h0 numpy arrays."""
    n = len(batch.q_len)
    out = np.zeros(n, dtype=EXT_DTYPE)
    ma, mp = _mat(mat)
    cells = C.c_int64(0)
    lib().orc_extend_batch(n, batch.arena.ctypes.data, batch.q_off.ctypes.data, batch.q_len.ctypes.data,
                           batch.t_off.ctypes.data, batch.t_len.ctypes.data, batch.h0.ctypes.data,
                           5, mp, o_del, e_del, o_ins, e_ins, w, end_bonus, zdrop,
                           out.ctypes.data, C.addressof(cells), nthreads)
    return (out, cells.value) if want_cells else out


def extend_batch_job_cells(batch, mat, o_del, e_del, o_ins, e_ins, w, end_bonus, zdrop, nthreads=8):
    """extend_batch plus the number of DP cells ksw_extend2 evaluates for each job."""
    n = len(batch.q_len)
    out = np.zeros(n, dtype=EXT_DTYPE)
    cells = np.zeros(n, dtype=np.int32)
    ma, mp = _mat(mat)
    lib().orc_extend_batch_cells(n, batch.arena.ctypes.data, batch.q_off.ctypes.data, batch.q_len.ctypes.data,
                                 batch.t_off.ctypes.data, batch.t_len.ctypes.data, batch.h0.ctypes.data,
                                 5, mp, o_del, e_del, o_ins, e_ins, w, end_bonus, zdrop,
                                 out.ctypes.data, cells.ctypes.data, nthreads)
    return out, cells


def ungapped_batch(batch, mat, o_del, e_del, o_ins, e_ins, zdrop):
    """The reference's ungapped fast path on every job: (results, accepted flags)."""
    n = len(batch.q_len)
    out = np.zeros(n, dtype=EXT_DTYPE)
    hit = np.zeros(n, dtype=np.uint8)
    ma, mp = _mat(mat)
    lib().orc_ungapped_batch(n, batch.arena.ctypes.data, batch.q_off.ctypes.data, batch.q_len.ctypes.data,
                             batch.t_off.ctypes.data, batch.t_len.ctypes.data, batch.h0.ctypes.data,
                             5, mp, o_del, e_del, o_ins, e_ins, zdrop, out.ctypes.data, hit.ctypes.data)
    return out, hit.astype(bool)


def ungapped_xmax(mat, o_del, e_del, o_ins, e_ins):
    ma, mp = _mat(mat)
    return lib().orc_ungapped_xmax(mp, o_del, e_del, o_ins, e_ins)


def align2_batch(batch, mat, o_del, e_del, o_ins, e_ins, minsc, max_sc, batched=True, nthreads=8,
                 want_cells=False):
    n = len(batch.q_len)
    out = np.zeros(n, dtype=KSW_DTYPE)
    ma, mp = _mat(mat)
    cells = C.c_int64(0)
    lib().orc_align2_batch(n, batch.arena.ctypes.data, batch.q_off.ctypes.data, batch.q_len.ctypes.data,
                           batch.t_off.ctypes.data, batch.t_len.ctypes.data, 5, mp,
                           o_del, e_del, o_ins, e_ins, minsc, max_sc, 1 if batched else 0,
                           out.ctypes.data, C.addressof(cells), nthreads)
    return (out, cells.value) if want_cells else out


def simd_extend_batch(batch, mat, o_del, e_del, o_ins, e_ins, w, end_bonus, zdrop, nthreads=8):
    """The SIMD CPU baseline (oracle/ksw_simd.c): the same results as extend_batch."""
    n = len(batch.q_len)
    out = np.zeros(n, dtype=EXT_DTYPE)
    ma, mp = _mat(mat)
    lib().orc_simd_extend_batch(n, batch.arena.ctypes.data, batch.q_off.ctypes.data, batch.q_len.ctypes.data,
                                batch.t_off.ctypes.data, batch.t_len.ctypes.data, batch.h0.ctypes.data,
                                5, mp, o_del, e_del, o_ins, e_ins, w, end_bonus, zdrop, out.ctypes.data, nthreads)
    return out


def simd_align2_batch(batch, mat, o_del, e_del, o_ins, e_ins, minsc, max_sc, nthreads=8):
    """The SIMD CPU baseline: the same results as align2_batch with batched semantics."""
    n = len(batch.q_len)
    out = np.zeros(n, dtype=KSW_DTYPE)
    ma, mp = _mat(mat)
    lib().orc_simd_align2_batch(n, batch.arena.ctypes.data, batch.q_off.ctypes.data, batch.q_len.ctypes.data,
                                batch.t_off.ctypes.data, batch.t_len.ctypes.data, 5, mp,
                                o_del, e_del, o_ins, e_ins, minsc, max_sc, out.ctypes.data, nthreads)
    return out


def _ptr_arrays(seqs):
    bufs = [np.ascontiguousarray(np.asarray(s, dtype=np.uint8)) for s in seqs]
    ptrs = (C.c_void_p * max(1, len(bufs)))(*[b.ctypes.data for b in bufs])
    lens = np.array([len(b) for b in bufs], dtype=np.int32)
    return bufs, ptrs, lens


def lockstep_extend(queries, targets, h0s, mat, o_del, e_del, o_ins, e_ins, w, end_bonus, zdrop,
                    lanes=16):
    qb, qp, ql = _ptr_arrays(queries)
    tb, tp, tl = _ptr_arrays(targets)
    h0 = np.asarray(h0s, dtype=np.int32)
    ma, mp = _mat(mat)
    out = (ExtResult * max(1, len(queries)))()
    lib().orc_lockstep_extend(len(queries), C.addressof(qp), ql.ctypes.data, C.addressof(tp),
                              tl.ctypes.data, h0.ctypes.data, 5, mp, o_del, e_del, o_ins, e_ins,
                              w, end_bonus, zdrop, lanes, C.addressof(out))
    return [out[k].tup() for k in range(len(queries))]


def lockstep_local_fwd(queries, targets, minsc, endsc, mat, o_del, e_del, o_ins, e_ins, max_sc,
                       lanes=16):
    qb, qp, ql = _ptr_arrays(queries)
    tb, tp, tl = _ptr_arrays(targets)
    mi = np.asarray(minsc, dtype=np.int32)
    en = np.asarray(endsc, dtype=np.int32)
    ma, mp = _mat(mat)
    out = (FwdResult * max(1, len(queries)))()
    lib().orc_lockstep_local_fwd(len(queries), C.addressof(qp), ql.ctypes.data, C.addressof(tp),
                                 tl.ctypes.data, mi.ctypes.data, en.ctypes.data, 5, mp,
                                 o_del, e_del, o_ins, e_ins, max_sc, lanes, C.addressof(out))
    return [out[k].tup() for k in range(len(queries))]


def brute_extend_unbanded(q, t, mat, o_del, e_del, o_ins, e_ins, h0):
    qa, qp = _buf(q)
    ta, tp = _buf(t)
    ma, mp = _mat(mat)
    return lib().orc_brute_extend_unbanded(len(qa), qp, len(ta), tp, 5, mp, o_del, e_del, o_ins,
                                           e_ins, h0)


def brute_gotoh_local(q, t, mat, o_del, e_del, o_ins, e_ins, pad_to=None):
    q = list(q)
    if pad_to is not None:
        q = q + [5] * (pad_to - len(q))
    qa, qp = _buf(q)
    ta, tp = _buf(t)
    ma, mp = _mat(mat)
    return lib().orc_brute_gotoh_local(len(qa), qp, len(ta), tp, 5, mp, o_del, e_del, o_ins, e_ins)


def brute_enum_local(q, t, mat, o_del, e_del, o_ins, e_ins):
    qa, qp = _buf(q)
    ta, tp = _buf(t)
    ma, mp = _mat(mat)
    s, te, qe = C.c_int32(), C.c_int32(), C.c_int32()
    lib().orc_brute_enum_local(len(qa), qp, len(ta), tp, 5, mp, o_del, e_del, o_ins, e_ins,
                               C.addressof(s), C.addressof(te), C.addressof(qe))
    return s.value, te.value, qe.value
