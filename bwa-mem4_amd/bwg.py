"""Python binding of libbwg.so, mirroring bwa-mem4's plug interface for this path.

``HipBackend`` mirrors the ``SwBackend`` trait (crates/bwa-extend/src/lib.rs:83-176): ``name``,
``extend`` and ``extend_batch`` with the same argument names, meaning and order.
``batched_ksw_align2`` mirrors the free function of crates/bwa-neon/src/matesw.rs:299-309.
``ExtendJob`` / ``KswJob`` / ``ExtendResult`` / ``KswAlignResult`` mirror the reference types
(crates/bwa-extend/src/lib.rs:181-195, crates/bwa-neon/src/matesw.rs:104-116,
crates/bwa-extend/src/sw.rs:138-162 and :749-774).

The reference's trait is infallible and a backend must never fall back silently; here every error
code from the library raises ``BwgError`` (a Rust shim maps it to ``panic!``).  Nothing in this
module computes a result on the CPU.
"""
import ctypes as C
import os
from collections import namedtuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbwg.so")
# BWG_WORKLOAD_SO selects another build of the host-only workload library (the sanitizer build)
WORKLOAD_LIB_PATH = os.environ.get("BWG_WORKLOAD_SO") or os.path.join(HERE, "libbwg_workload.so")

ExtendResult = namedtuple("ExtendResult", "score qle tle gtle gscore max_off")
KswAlignResult = namedtuple("KswAlignResult", "score qb qe tb te score2 te2")
ExtendJob = namedtuple("ExtendJob", "query target h0")
KswJob = namedtuple("KswJob", "query target")

EXT_DTYPE = np.dtype([(n, "<i4") for n in ExtendResult._fields])
KSW_DTYPE = np.dtype([(n, "<i4") for n in KswAlignResult._fields])

BWG_OK, BWG_EIO, BWG_E2BIG, BWG_EAGAIN, BWG_ENOMEM, BWG_ENODEV, BWG_EINVAL = 0, -5, -7, -11, -12, -19, -22
FORM_PLAIN, FORM_BATCHED = 0, 1
OPT_FORM, OPT_CONCURRENT_BINS, OPT_GRAPHS, OPT_EXT_THREAD_MAX_Q = 1, 2, 3, 4
OPT_AUTOTUNE, OPT_EXT_K_STEP, OPT_KSW_K_STEP, OPT_EXT_SCORE_WIDTH, OPT_UNGAPPED = 5, 6, 7, 8, 9
OPT_KSW_SCORE_WIDTH, OPT_LANES, OPT_POOL_CACHE_MB, OPT_POOL_PEAK_RESET = 10, 11, 12, 13
OPT_ITEM_TIMING, OPT_COALESCE = 14, 15
OPT_HOST_THREADS = 16
OPT_EXT_GROUP_LANES = 17
OPT_KSW_GROUP_LANES = 18
OPT_BIN_STREAMS = 19
STAT_LAUNCHES, STAT_JOBS, STAT_LAST_KERNEL_NS, STAT_LAST_MAIN_NS, STAT_LAST_MAIN_LAUNCHES = 100, 101, 102, 103, 104
STAT_POOL_IN_USE, STAT_POOL_CACHED, STAT_POOL_PEAK = 105, 106, 107
# host wall time of the phases of the last synchronous extension / rescue call (ns)
STAT_PHASES = {"validate": 110, "stage": 111, "plan": 112, "kernels": 113, "fetch": 114, "total": 115}
MAX_JOB_LEN = 1 << 24

# Every symbol include/bwg.h declares.
EXPORTED_SYMBOLS = (
    "bwg_create", "bwg_destroy", "bwg_name", "bwg_strerror", "bwg_extend_batch", "bwg_extend",
    "bwg_ksw_align2_batch", "bwg_set_option", "bwg_get_stat", "bwg_batch_stage_ext",
    "bwg_batch_stage_ksw", "bwg_batch_run", "bwg_batch_results", "bwg_batch_destroy",
    "bwg_server_create", "bwg_server_set_reference", "bwg_server_submit", "bwg_server_poll",
    "bwg_server_wait", "bwg_server_stats", "bwg_server_destroy", "bwg_hbm_copy", "bwg_tune_report",
    "bwg_batch_ungapped_hits", "bwg_extend_side_batch", "bwg_batch_stage_side", "bwg_batch_side_bands",
    "bwg_set_reference", "bwg_matesw_batch", "bwg_batch_stage_matesw", "bwg_batch_windows",
    "bwg_global_batch", "bwg_batch_stage_global", "bwg_batch_cigar_lengths", "bwg_batch_cigars",
    "bwg_extend_batch_noalloc", "bwg_batch_report", "bwg_batch_job_bins",
)

# bwg_mate_anchor, bwg_rescue_window, bwg_pe_orient and bwg_contig as numpy records
ANCHOR_DTYPE = np.dtype([("rb", "<i8"), ("rid", "<i4"), ("l_ms", "<u4"), ("ms_off", "<u8"),
                         ("ma_begin", "<u8"), ("ma_end", "<u8")])
WINDOW_DTYPE = np.dtype([("rb", "<i8"), ("re", "<i8"), ("rid", "<i4"), ("flags", "<i4")])
PE_DTYPE = np.dtype([("low", "<i4"), ("high", "<i4"), ("failed", "<i4")])
CONTIG_DTYPE = np.dtype([("offset", "<i8"), ("len", "<i4"), ("_pad", "<i4")])
WIN_VALID, WIN_REV, WIN_SKIPPED = 1, 2, 4


class BwgError(RuntimeError):
    def __init__(self, code, what):
        self.code = code
        super().__init__("%s: %s (%d)" % (what, _lib().bwg_strerror(code).decode(), code))


class ExtParams(C.Structure):
    _fields_ = [("m", C.c_int32), ("mat", C.c_int8 * 25), ("o_del", C.c_int32),
                ("e_del", C.c_int32), ("o_ins", C.c_int32), ("e_ins", C.c_int32),
                ("w", C.c_int32), ("end_bonus", C.c_int32), ("zdrop", C.c_int32)]


class KswParams(C.Structure):
    _fields_ = [("m", C.c_int32), ("mat", C.c_int8 * 25), ("o_del", C.c_int32),
                ("e_del", C.c_int32), ("o_ins", C.c_int32), ("e_ins", C.c_int32),
                ("minsc", C.c_int32), ("max_sc", C.c_int32)]


class GlobalParams(C.Structure):
    _fields_ = [("m", C.c_int32), ("mat", C.c_int8 * 25), ("o_del", C.c_int32),
                ("e_del", C.c_int32), ("o_ins", C.c_int32), ("e_ins", C.c_int32)]


_L = None
_P = C.c_void_p


def _lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libbwg.so is not built (run __graft_entry__.build()): %s" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        L.bwg_create.restype = C.c_int
        L.bwg_create.argtypes = [C.c_int, _P]
        L.bwg_destroy.restype = None
        L.bwg_destroy.argtypes = [_P]
        L.bwg_name.restype = C.c_char_p
        L.bwg_name.argtypes = [_P]
        L.bwg_strerror.restype = C.c_char_p
        L.bwg_strerror.argtypes = [C.c_int]
        L.bwg_extend_batch.restype = C.c_int
        L.bwg_extend_batch.argtypes = [_P, _P, C.c_size_t, _P, C.c_size_t, _P, _P, _P, _P, _P, _P]
        L.bwg_batch_report.restype = C.c_int
        L.bwg_batch_report.argtypes = [_P, _P, C.c_char_p, C.c_size_t]
        L.bwg_batch_job_bins.restype = C.c_int
        L.bwg_batch_job_bins.argtypes = [_P, _P, _P]
        L.bwg_extend_batch_noalloc.restype = C.c_int
        L.bwg_extend_batch_noalloc.argtypes = [_P, _P, C.c_size_t, _P, C.c_size_t, _P, _P, _P, _P, _P, _P]
        L.bwg_extend.restype = C.c_int
        L.bwg_extend.argtypes = [_P, _P, _P, C.c_uint32, _P, C.c_uint32, C.c_int32, _P]
        L.bwg_ksw_align2_batch.restype = C.c_int
        L.bwg_ksw_align2_batch.argtypes = [_P, _P, C.c_size_t, _P, C.c_size_t, _P, _P, _P, _P, _P]
        L.bwg_set_option.restype = C.c_int
        L.bwg_set_option.argtypes = [_P, C.c_int, C.c_int64]
        L.bwg_get_stat.restype = C.c_int64
        L.bwg_get_stat.argtypes = [_P, C.c_int]
        L.bwg_batch_stage_ext.restype = C.c_int
        L.bwg_batch_stage_ext.argtypes = [_P, _P, C.c_size_t, _P, C.c_size_t, _P, _P, _P, _P, _P, _P]
        L.bwg_batch_stage_ksw.restype = C.c_int
        L.bwg_batch_stage_ksw.argtypes = [_P, _P, C.c_size_t, _P, C.c_size_t, _P, _P, _P, _P, _P]
        L.bwg_batch_run.restype = C.c_int
        L.bwg_batch_run.argtypes = [_P, _P]
        L.bwg_batch_results.restype = C.c_int
        L.bwg_batch_results.argtypes = [_P, _P, _P]
        L.bwg_batch_destroy.restype = None
        L.bwg_batch_destroy.argtypes = [_P, _P]
        L.bwg_hbm_copy.restype = C.c_int
        L.bwg_hbm_copy.argtypes = [_P, C.c_size_t, C.c_int, _P]
        L.bwg_batch_ungapped_hits.restype = C.c_int
        L.bwg_batch_ungapped_hits.argtypes = [_P, _P, _P]
        L.bwg_tune_report.restype = C.c_int
        L.bwg_tune_report.argtypes = [_P, C.c_char_p, C.c_size_t]
        L.bwg_extend_side_batch.restype = C.c_int
        L.bwg_extend_side_batch.argtypes = [_P, _P, C.c_int, C.c_size_t, _P, C.c_size_t, _P, _P, _P, _P, _P, _P,
                                            _P, _P]
        L.bwg_batch_stage_side.restype = C.c_int
        L.bwg_batch_stage_side.argtypes = [_P, _P, C.c_int, C.c_size_t, _P, C.c_size_t, _P, _P, _P, _P, _P, _P,
                                           _P]
        L.bwg_batch_side_bands.restype = C.c_int
        L.bwg_batch_side_bands.argtypes = [_P, _P, _P]
        L.bwg_set_reference.restype = C.c_int
        L.bwg_set_reference.argtypes = [_P, _P, C.c_int64, _P, C.c_int32]
        L.bwg_matesw_batch.restype = C.c_int
        L.bwg_matesw_batch.argtypes = [_P, _P, C.c_int32, _P, C.c_size_t, _P, _P, C.c_size_t, _P, C.c_size_t,
                                       _P, _P]
        L.bwg_batch_stage_matesw.restype = C.c_int
        L.bwg_batch_stage_matesw.argtypes = [_P, _P, C.c_int32, _P, C.c_size_t, _P, _P, C.c_size_t, _P,
                                             C.c_size_t, _P]
        L.bwg_batch_windows.restype = C.c_int
        L.bwg_batch_windows.argtypes = [_P, _P, _P]
        L.bwg_global_batch.restype = C.c_int
        L.bwg_global_batch.argtypes = [_P, _P, C.c_size_t, _P, C.c_size_t, _P, _P, _P, _P, _P, _P, _P, _P,
                                       C.c_size_t]
        L.bwg_batch_stage_global.restype = C.c_int
        L.bwg_batch_stage_global.argtypes = [_P, _P, C.c_size_t, _P, C.c_size_t, _P, _P, _P, _P, _P, _P]
        L.bwg_batch_cigar_lengths.restype = C.c_int
        L.bwg_batch_cigar_lengths.argtypes = [_P, _P, _P]
        L.bwg_batch_cigars.restype = C.c_int
        L.bwg_batch_cigars.argtypes = [_P, _P, _P, C.c_size_t]
        _L = L
    return _L


def library():
    """The loaded ctypes handle of libbwg.so."""
    return _lib()


def _check(rc, what):
    if rc != BWG_OK:
        raise BwgError(rc, what)


def _mat25(mat):
    a = np.asarray(mat, dtype=np.int8).reshape(-1)
    if a.size != 25:
        raise ValueError("mat must have 25 entries")
    return (C.c_int8 * 25)(*[int(x) for x in a])


def ext_params(m, mat, o_del, e_del, o_ins, e_ins, w, end_bonus, zdrop):
    return ExtParams(m, _mat25(mat), o_del, e_del, o_ins, e_ins, w, end_bonus, zdrop)


def global_params(m, mat, o_del, e_del, o_ins, e_ins):
    return GlobalParams(m, _mat25(mat), o_del, e_del, o_ins, e_ins)


def ksw_params(m, mat, o_del, e_del, o_ins, e_ins, minsc, max_sc):
    return KswParams(m, _mat25(mat), o_del, e_del, o_ins, e_ins, minsc, max_sc)


class JobBatch:
    """Jobs as one byte arena plus per-job offsets/lengths (the layout the C ABI takes)."""

    def __init__(self, arena, q_off, q_len, t_off, t_len, h0=None, kind=None):
        self.arena = np.ascontiguousarray(arena, dtype=np.uint8)
        self.q_off = np.ascontiguousarray(q_off, dtype=np.uint64)
        self.q_len = np.ascontiguousarray(q_len, dtype=np.uint32)
        self.t_off = np.ascontiguousarray(t_off, dtype=np.uint64)
        self.t_len = np.ascontiguousarray(t_len, dtype=np.uint32)
        n = len(self.q_len)
        self.h0 = np.ascontiguousarray(h0 if h0 is not None else np.zeros(n), dtype=np.int32)
        self.kind = None if kind is None else np.ascontiguousarray(kind, dtype=np.uint8)

    def __len__(self):
        return len(self.q_len)

    @classmethod
    def from_pairs(cls, queries, targets, h0=None):
        parts, q_off, q_len, t_off, t_len = [], [], [], [], []
        o = 0
        for q, t in zip(queries, targets):
            q = np.asarray(q, dtype=np.uint8)
            t = np.asarray(t, dtype=np.uint8)
            q_off.append(o)
            q_len.append(len(q))
            parts.append(q)
            o += len(q)
            t_off.append(o)
            t_len.append(len(t))
            parts.append(t)
            o += len(t)
        arena = np.concatenate(parts) if parts else np.zeros(0, dtype=np.uint8)
        return cls(arena, q_off, q_len, t_off, t_len, h0)

    def subset(self, idx):
        """A batch of the jobs `idx` (re-packed into a fresh arena)."""
        idx = np.asarray(idx)
        qs = [self.arena[int(self.q_off[k]):int(self.q_off[k]) + int(self.q_len[k])] for k in idx]
        ts = [self.arena[int(self.t_off[k]):int(self.t_off[k]) + int(self.t_len[k])] for k in idx]
        return JobBatch.from_pairs(qs, ts, self.h0[idx])

    def query(self, k):
        return self.arena[int(self.q_off[k]):int(self.q_off[k]) + int(self.q_len[k])]

    def target(self, k):
        return self.arena[int(self.t_off[k]):int(self.t_off[k]) + int(self.t_len[k])]

    def _args(self):
        return (len(self.q_len), self.arena.ctypes.data, self.arena.size, self.q_off.ctypes.data,
                self.q_len.ctypes.data, self.t_off.ctypes.data, self.t_len.ctypes.data)


class Context:
    """One libbwg context (a HIP stream and device buffers on one GPU)."""

    def __init__(self, device=0, form=None):
        h = C.c_void_p()
        _check(_lib().bwg_create(device, C.byref(h)), "bwg_create(device=%d)" % device)
        self.h = h
        if form is not None:
            self.set_form(form)

    def close(self):
        if self.h:
            _lib().bwg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_form(self, form):
        _check(_lib().bwg_set_option(self.h, OPT_FORM, int(form)), "bwg_set_option(form)")

    def set_option(self, key, value):
        _check(_lib().bwg_set_option(self.h, int(key), int(value)), "bwg_set_option(%d)" % key)

    def stat(self, key):
        return int(_lib().bwg_get_stat(self.h, key))

    def phases(self):
        """Host wall time (seconds) of the phases of the last synchronous extension or rescue call."""
        return {k: self.stat(v) * 1e-9 for k, v in STAT_PHASES.items()}

    def tune_report(self):
        """The autotuner's table: one line per tuned length class, chosen mapping starred."""
        n = _lib().bwg_tune_report(self.h, None, 0)
        _check(min(n, 0), "bwg_tune_report")
        buf = C.create_string_buffer(n + 1)
        _lib().bwg_tune_report(self.h, buf, n + 1)
        return buf.value.decode()

    def hbm_copy(self, nbytes, reps=1):
        """Copy `nbytes` between two device buffers `reps` times with the library's streaming copy
        kernel; returns the rate in GB/s counting bytes read plus bytes written."""
        gbs = C.c_double(0.0)
        _check(_lib().bwg_hbm_copy(self.h, int(nbytes), int(reps), C.byref(gbs)), "bwg_hbm_copy")
        return gbs.value

    def name(self):
        return _lib().bwg_name(self.h).decode()

    def extend_arrays(self, batch, params):
        out = np.zeros(len(batch), dtype=EXT_DTYPE)
        n, ap, al, qo, ql, to, tl = batch._args()
        _check(_lib().bwg_extend_batch(self.h, C.byref(params), n, ap, al, qo, ql, to, tl,
                                       batch.h0.ctypes.data, out.ctypes.data), "bwg_extend_batch")
        return out

    def align2_arrays(self, batch, params):
        out = np.zeros(len(batch), dtype=KSW_DTYPE)
        n, ap, al, qo, ql, to, tl = batch._args()
        _check(_lib().bwg_ksw_align2_batch(self.h, C.byref(params), n, ap, al, qo, ql, to, tl,
                                           out.ctypes.data), "bwg_ksw_align2_batch")
        return out

    def extend_side_arrays(self, batch, params, prev, max_band_try=2):
        """One side of the seed extensions with bwa's band doubling (the reference's run_side,
        crates/bwa-mem/src/across.rs:1357-1460): (results, accepted band per job)."""
        out = np.zeros(len(batch), dtype=EXT_DTYPE)
        w_out = np.zeros(len(batch), dtype=np.int32)
        prev = np.ascontiguousarray(prev, dtype=np.int32)
        n, ap, al, qo, ql, to, tl = batch._args()
        _check(_lib().bwg_extend_side_batch(self.h, C.byref(params), int(max_band_try), n, ap, al, qo, ql, to, tl,
                                            batch.h0.ctypes.data, prev.ctypes.data, out.ctypes.data,
                                            w_out.ctypes.data), "bwg_extend_side_batch")
        return out, w_out

    def global_arrays(self, batch, params, w):
        """ksw_global2 on every job with its own band w[k]: (scores, runs per job, the runs of all
        jobs concatenated as len << 4 | op words)."""
        n = len(batch)
        w = np.ascontiguousarray(w, dtype=np.int32)
        score = np.zeros(n, dtype=np.int32)
        ncig = np.zeros(n, dtype=np.uint32)
        cap = int(np.sum(batch.q_len.astype(np.int64) + batch.t_len + 2)) if n else 0
        cig = np.zeros(max(1, cap), dtype=np.uint32)
        nn, ap, al, qo, ql, to, tl = batch._args()
        _check(_lib().bwg_global_batch(self.h, C.byref(params), nn, ap, al, qo, ql, to, tl, w.ctypes.data,
                                       score.ctypes.data, ncig.ctypes.data, cig.ctypes.data, cap),
               "bwg_global_batch")
        return score, ncig, cig[:int(ncig.sum())]

    def stage_global(self, batch, params, w):
        h = C.c_void_p()
        w = np.ascontiguousarray(w, dtype=np.int32)
        n, ap, al, qo, ql, to, tl = batch._args()
        _check(_lib().bwg_batch_stage_global(self.h, C.byref(params), n, ap, al, qo, ql, to, tl, w.ctypes.data,
                                             C.byref(h)), "bwg_batch_stage_global")
        sb = StagedBatch(self, h, n, np.int32)
        sb.cigar_cap = int(np.sum(batch.q_len.astype(np.int64) + batch.t_len + 2)) if n else 0
        return sb

    def set_reference(self, fwd, contigs):
        """Upload a reference (forward strand codes, one byte per base) and its contigs
        (CONTIG_DTYPE records, ascending offsets) for mate rescue."""
        fwd = np.ascontiguousarray(fwd, dtype=np.uint8)
        contigs = np.ascontiguousarray(contigs, dtype=CONTIG_DTYPE)
        _check(_lib().bwg_set_reference(self.h, fwd.ctypes.data, len(fwd), contigs.ctypes.data, len(contigs)),
               "bwg_set_reference")

    def matesw_arrays(self, rnd, params):
        """One mate-rescue round (RescueRound) on the device: (windows, results), 4 per anchor."""
        n = len(rnd.anchors)
        win = np.zeros(4 * n, dtype=WINDOW_DTYPE)
        out = np.zeros(4 * n, dtype=KSW_DTYPE)
        _check(_lib().bwg_matesw_batch(self.h, C.byref(params), int(rnd.min_seed_len), rnd.pes.ctypes.data, n,
                                       rnd.anchors.ctypes.data, rnd.ma_rb.ctypes.data, rnd.n_ma,
                                       rnd.arena.ctypes.data, len(rnd.arena), win.ctypes.data, out.ctypes.data),
               "bwg_matesw_batch")
        return win, out

    def stage_matesw(self, rnd, params):
        h = C.c_void_p()
        n = len(rnd.anchors)
        _check(_lib().bwg_batch_stage_matesw(self.h, C.byref(params), int(rnd.min_seed_len), rnd.pes.ctypes.data,
                                             n, rnd.anchors.ctypes.data, rnd.ma_rb.ctypes.data, rnd.n_ma,
                                             rnd.arena.ctypes.data, len(rnd.arena), C.byref(h)),
               "bwg_batch_stage_matesw")
        return StagedBatch(self, h, 4 * n, KSW_DTYPE)

    def stage_side(self, batch, params, prev, max_band_try=2):
        """Stage a band-retry side batch; run it with StagedBatch.run, bands with .side_bands()."""
        h = C.c_void_p()
        prev = np.ascontiguousarray(prev, dtype=np.int32)
        n, ap, al, qo, ql, to, tl = batch._args()
        _check(_lib().bwg_batch_stage_side(self.h, C.byref(params), int(max_band_try), n, ap, al, qo, ql, to, tl,
                                           batch.h0.ctypes.data, prev.ctypes.data, C.byref(h)),
               "bwg_batch_stage_side")
        return StagedBatch(self, h, n, EXT_DTYPE)

    def stage(self, batch, params):
        """Stage a batch into device memory; returns a StagedBatch that can be run repeatedly."""
        h = C.c_void_p()
        n, ap, al, qo, ql, to, tl = batch._args()
        if isinstance(params, ExtParams):
            _check(_lib().bwg_batch_stage_ext(self.h, C.byref(params), n, ap, al, qo, ql, to, tl,
                                              batch.h0.ctypes.data, C.byref(h)), "bwg_batch_stage_ext")
            return StagedBatch(self, h, n, EXT_DTYPE)
        _check(_lib().bwg_batch_stage_ksw(self.h, C.byref(params), n, ap, al, qo, ql, to, tl,
                                          C.byref(h)), "bwg_batch_stage_ksw")
        return StagedBatch(self, h, n, KSW_DTYPE)


class StagedBatch:
    def __init__(self, ctx, h, n, dtype):
        self.ctx, self.h, self.n, self.dtype = ctx, h, n, dtype

    def run(self):
        _check(_lib().bwg_batch_run(self.ctx.h, self.h), "bwg_batch_run")

    def results(self):
        out = np.zeros(self.n, dtype=self.dtype)
        _check(_lib().bwg_batch_results(self.ctx.h, self.h, out.ctypes.data), "bwg_batch_results")
        return out

    def cigars(self):
        """Runs per job and the concatenated runs of a global-alignment batch's last run."""
        ncig = np.zeros(self.n, dtype=np.uint32)
        _check(_lib().bwg_batch_cigar_lengths(self.ctx.h, self.h, ncig.ctypes.data), "bwg_batch_cigar_lengths")
        cig = np.zeros(max(1, int(ncig.sum())), dtype=np.uint32)
        _check(_lib().bwg_batch_cigars(self.ctx.h, self.h, cig.ctypes.data, len(cig)), "bwg_batch_cigars")
        return ncig, cig[:int(ncig.sum())]

    def windows(self):
        """Rescue windows of a mate-rescue batch's last run (4 per anchor)."""
        w = np.zeros(self.n, dtype=WINDOW_DTYPE)
        _check(_lib().bwg_batch_windows(self.ctx.h, self.h, w.ctypes.data), "bwg_batch_windows")
        return w

    def side_bands(self):
        """Accepted band of every job of a side batch's last run."""
        w = np.zeros(self.n, dtype=np.int32)
        _check(_lib().bwg_batch_side_bands(self.ctx.h, self.h, w.ctypes.data), "bwg_batch_side_bands")
        return w

    def report(self):
        """The launches of the last batched run: dicts with index, round, stage, kernel, jobs, grid
        and us (device time; -1 unless OPT_ITEM_TIMING was on)."""
        n = _lib().bwg_batch_report(self.ctx.h, self.h, None, 0)
        _check(min(n, 0), "bwg_batch_report")
        buf = C.create_string_buffer(n + 1)
        _lib().bwg_batch_report(self.ctx.h, self.h, buf, n + 1)
        rows = []
        for line in buf.value.decode().splitlines():
            f = line.split()
            d = {"index": int(f[0]), "kernel": f[3]}
            for kv in f[1:3] + f[4:]:
                k, v = kv.split("=")
                d[k] = float(v) if k == "us" else int(v)
            rows.append(d)
        return rows

    def job_bins(self):
        """Each job's launch index (into report()) of the bin it was planned into, -1 for none."""
        out = np.zeros(self.n, dtype=np.int32)
        _check(_lib().bwg_batch_job_bins(self.ctx.h, self.h, out.ctypes.data), "bwg_batch_job_bins")
        return out

    def ungapped_hits(self):
        """Jobs of the last batched run resolved by the ungapped-diagonal fast path."""
        v = C.c_uint64(0)
        _check(_lib().bwg_batch_ungapped_hits(self.ctx.h, self.h, C.byref(v)), "bwg_batch_ungapped_hits")
        return int(v.value)

    def close(self):
        if self.h:
            _lib().bwg_batch_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HipBackend:
    """The SwBackend mirror (crates/bwa-extend/src/lib.rs:83-176) on one MI355X."""

    def __init__(self, device=0, form=None):
        self.ctx = Context(device, form)

    def name(self):
        return self.ctx.name()

    def extend(self, query, target, m, mat, o_del, e_del, o_ins, e_ins, w, end_bonus, zdrop, h0):
        q = np.ascontiguousarray(query, dtype=np.uint8)
        t = np.ascontiguousarray(target, dtype=np.uint8)
        out = np.zeros(1, dtype=EXT_DTYPE)
        p = ext_params(m, mat, o_del, e_del, o_ins, e_ins, w, end_bonus, zdrop)
        _check(_lib().bwg_extend(self.ctx.h, C.byref(p), q.ctypes.data, q.size, t.ctypes.data,
                                 t.size, int(h0), out.ctypes.data), "bwg_extend")
        return ExtendResult(*[int(x) for x in out[0]])

    def extend_batch(self, jobs, m, mat, o_del, e_del, o_ins, e_ins, w, end_bonus, zdrop):
        if len(jobs) == 0:
            return []
        batch = JobBatch.from_pairs([j.query for j in jobs], [j.target for j in jobs],
                                    [j.h0 for j in jobs])
        p = ext_params(m, mat, o_del, e_del, o_ins, e_ins, w, end_bonus, zdrop)
        out = self.ctx.extend_arrays(batch, p)
        return [ExtendResult(*[int(x) for x in r]) for r in out]

    def batched_ksw_align2(self, jobs, m, mat, o_del, e_del, o_ins, e_ins, minsc, max_sc):
        if len(jobs) == 0:
            return []
        batch = JobBatch.from_pairs([j.query for j in jobs], [j.target for j in jobs])
        p = ksw_params(m, mat, o_del, e_del, o_ins, e_ins, minsc, max_sc)
        out = self.ctx.align2_arrays(batch, p)
        return [KswAlignResult(*[int(x) for x in r]) for r in out]


_default_backend = None


def batched_ksw_align2(jobs, m, mat, o_del, e_del, o_ins, e_ins, minsc, max_sc):
    """Free-function form (crates/bwa-neon/src/matesw.rs:299-309) on device 0."""
    global _default_backend
    if _default_backend is None:
        _default_backend = HipBackend(0)
    return _default_backend.batched_ksw_align2(jobs, m, mat, o_del, e_del, o_ins, e_ins, minsc, max_sc)


# ---- synthetic workloads ------------------------------------------------------------------------
_W = None


def _wlib():
    global _W
    if _W is None:
        W = C.CDLL(WORKLOAD_LIB_PATH)
        W.bwgw_generate.restype = C.c_void_p
        W.bwgw_generate.argtypes = [C.c_int, C.c_uint64, C.c_uint64]
        W.bwgw_arena_len.restype = C.c_uint64
        W.bwgw_arena_len.argtypes = [C.c_void_p]
        W.bwgw_count.restype = C.c_uint64
        W.bwgw_count.argtypes = [C.c_void_p]
        W.bwgw_copy.restype = None
        W.bwgw_copy.argtypes = [C.c_void_p] + [C.c_void_p] * 7
        W.bwgw_free.restype = None
        W.bwgw_free.argtypes = [C.c_void_p]
        _W = W
    return _W


WORKLOAD_EXT_SMALL, WORKLOAD_RESCUE, WORKLOAD_EXT_TAIL, WORKLOAD_SERVE_MIX = 1, 2, 3, 4


def generate(kind, n, seed=42):
    """Synthetic jobs: 1 = extension, qlen U[1,130] from 150 bp reads, tlen qlen+100, h0
    U[19,150]; 2 = rescue, 150 bp mate vs a U[600,1200] window, 30% unplaced; 3 = long-tail
    extension (reads 100/150/250 bp, qlen 70% <30 / 25% 30-150 / 5% 150-250, h0 U[19,250]);
    4 = serving mix (80% kind 3, 20% kind 2)."""
    W = _wlib()
    h = W.bwgw_generate(kind, seed, n)
    try:
        n = int(W.bwgw_count(h))
        arena = np.zeros(int(W.bwgw_arena_len(h)) + 1, dtype=np.uint8)
        q_off = np.zeros(n, dtype=np.uint64)
        q_len = np.zeros(n, dtype=np.uint32)
        t_off = np.zeros(n, dtype=np.uint64)
        t_len = np.zeros(n, dtype=np.uint32)
        h0 = np.zeros(n, dtype=np.int32)
        kinds = np.zeros(n, dtype=np.uint8)
        W.bwgw_copy(h, arena.ctypes.data, q_off.ctypes.data, q_len.ctypes.data, t_off.ctypes.data,
                    t_len.ctypes.data, h0.ctypes.data, kinds.ctypes.data)
    finally:
        W.bwgw_free(h)
    return JobBatch(arena, q_off, q_len, t_off, t_len, h0, kinds)


class RescueRound:
    """One mate-rescue round's inputs: the reference (forward codes and contigs), insert windows per
    orientation, and anchors whose mates sit in `arena`, with the mates' existing regions ma_rb."""

    def __init__(self, fwd, contigs, pes, anchors, ma_rb, arena, min_seed_len):
        self.fwd = np.ascontiguousarray(fwd, dtype=np.uint8)
        self.contigs = np.ascontiguousarray(contigs, dtype=CONTIG_DTYPE)
        self.pes = np.ascontiguousarray(pes, dtype=PE_DTYPE)
        self.anchors = np.ascontiguousarray(anchors, dtype=ANCHOR_DTYPE)
        self.ma_rb = np.ascontiguousarray(ma_rb if len(ma_rb) else np.zeros(1), dtype=np.int64)[:max(1, len(ma_rb))]
        self.n_ma = len(ma_rb)
        self.arena = np.ascontiguousarray(arena, dtype=np.uint8)
        self.min_seed_len = int(min_seed_len)

    @property
    def l_pac(self):
        return len(self.fwd)


def _revcomp(s):
    s = np.asarray(s, dtype=np.uint8)[::-1]
    return np.where(s < 4, 3 - s, s).astype(np.uint8)


def generate_rescue_round(n_pairs, genome_len=2_000_000, n_contigs=4, read_len=150, seed=42,
                          insert_mean=550, insert_sd=150, random_frac=0.3, sub_rate=0.01):
    """A synthetic rescue round: a random genome cut into contigs, read pairs drawn as FR fragments
    (1% substitutions), and one anchor per pair, read 1's region on the forward strand or read 2's
    on the reverse half, with the other read as the mate; `random_frac` of the mates are random
    sequence.  Orientation windows follow bwa's FR-dominated statistics (RF and FF/RR failed except
    for some rounds that enable them); a few anchors carry a mate region already at a proper
    distance, which makes matesw_collect skip that orientation."""
    rng = np.random.default_rng(seed)
    fwd = rng.integers(0, 4, genome_len).astype(np.uint8)
    cuts = np.sort(rng.choice(np.arange(1, genome_len), n_contigs - 1, replace=False))
    starts = np.concatenate([[0], cuts])
    lens = np.diff(np.concatenate([starts, [genome_len]]))
    contigs = np.zeros(n_contigs, dtype=CONTIG_DTYPE)
    contigs["offset"] = starts
    contigs["len"] = lens
    l2 = 2 * genome_len
    pes = np.zeros(4, dtype=PE_DTYPE)
    lo, hi = max(1, insert_mean - 4 * insert_sd), insert_mean + 4 * insert_sd
    pes[:] = (lo, hi, 1)
    pes[1] = (lo, hi, 0)                      # FR
    if seed % 2:
        pes[2] = (max(1, lo // 2), hi // 2, 0)  # RF too, in some rounds
    n, l = n_pairs, read_len
    c = rng.integers(0, n_contigs, n)
    isize = np.clip(np.rint(rng.normal(insert_mean, insert_sd, n)), l, 4 * insert_mean).astype(np.int64)
    c = np.where(lens[c] <= isize + 2, int(np.argmax(lens)), c)
    x = starts[c] + (rng.random(n) * np.maximum(1, lens[c] - isize)).astype(np.int64)
    cols = np.arange(l)
    r1 = fwd[x[:, None] + cols]
    g2 = fwd[(x + isize - l)[:, None] + cols][:, ::-1]
    r2 = np.where(g2 < 4, 3 - g2, g2).astype(np.uint8)
    for r in (r1, r2):
        m = rng.random(r.shape) < sub_rate
        r[m] = (r[m] + rng.integers(1, 4, int(m.sum()))) % 4
    first = rng.random(n) < 0.5  # anchor read 1 (forward strand), else read 2 on the reverse half
    rb = np.where(first, x, l2 - (x + isize))
    mates = np.where(first[:, None], r2, r1)
    unplaced = rng.random(n) < random_frac
    mates[unplaced] = rng.integers(0, 4, (int(unplaced.sum()), l))
    # existing mate regions: one at a proper FR distance (5%), one unrelated (5%)
    paired = rng.random(n) < 0.05
    other = rng.random(n) < 0.05
    cnt = paired.astype(np.int64) + other
    ma_end = np.cumsum(cnt)
    ma_begin = ma_end - cnt
    ma_rb = np.zeros(int(cnt.sum()), dtype=np.int64)
    ma_rb[ma_begin[paired]] = np.where(first, l2 - (x + isize), x)[paired]
    ma_rb[(ma_begin + paired)[other]] = rng.integers(0, l2, int(other.sum()))
    anchors = np.zeros(n, dtype=ANCHOR_DTYPE)
    anchors["rb"] = rb
    anchors["rid"] = c
    anchors["l_ms"] = l
    anchors["ms_off"] = np.arange(n, dtype=np.uint64) * l
    anchors["ma_begin"] = ma_begin
    anchors["ma_end"] = ma_end
    arena = np.ascontiguousarray(mates.reshape(-1), dtype=np.uint8)
    return RescueRound(fwd, contigs, pes, anchors, np.asarray(ma_rb, dtype=np.int64), arena, 19)


def global_band(qlen, tlen, a=1, o_del=6, e_del=1, o_ins=6, e_ins=1, w_region=100):
    """The band the reference's region-to-CIGAR step gives ksw_global2 (crates/bwa-mem/src/
    cigar.rs:417-444): the affordable gap for half the read, plus the length difference, capped
    by the region's band and floored at the length difference + 3."""
    half = ((np.asarray(qlen, np.int64) + 1) >> 1) * a
    max_ins = ((half - o_ins) / e_ins + 1.0).astype(np.int64)
    max_del = ((half - o_del) / e_del + 1.0).astype(np.int64)
    max_gap = np.maximum(np.maximum(max_ins, max_del), 1)
    diff = np.abs(np.asarray(tlen, np.int64) - np.asarray(qlen, np.int64))
    w = (max_gap + diff + 1) >> 1
    w = np.minimum(w, w_region)
    return np.maximum(w, diff + 3).astype(np.int32)


def generate_global_jobs(n, seed=42, read_lens=(100, 150, 250), sub_rate=0.01, indel_frac=0.3, max_indel=12):
    """Synthetic region-to-CIGAR jobs: a read segment as query and its reference span as target
    (1% substitutions; `indel_frac` of the jobs with one or two indels of 1..max_indel bases),
    each with the band global_band() gives.  Returns (JobBatch, w)."""
    rng = np.random.default_rng(seed)
    lens = rng.choice(np.asarray(read_lens), n)
    qs, ts = [], []
    for k in range(n):
        L = int(lens[k])
        q = rng.integers(0, 4, L).astype(np.uint8)
        t = q.copy()
        m = rng.random(L) < sub_rate
        t[m] = (t[m] + rng.integers(1, 4, int(m.sum()))) % 4
        if rng.random() < indel_frac:
            for _ in range(int(rng.integers(1, 3))):
                p = int(rng.integers(1, len(t)))
                g = int(rng.integers(1, max_indel + 1))
                if rng.random() < 0.5:
                    t = np.concatenate([t[:p], rng.integers(0, 4, g).astype(np.uint8), t[p:]])
                else:
                    t = np.concatenate([t[:p], t[min(len(t), p + g):]])
        if len(t) == 0:
            t = np.zeros(1, np.uint8)
        qs.append(q)
        ts.append(t)
    batch = JobBatch.from_pairs(qs, ts)
    return batch, global_band(batch.q_len, batch.t_len)


def default_mat(a=1, b=4):
    """bwa's scoring matrix (crates/bwa-core/src/opt.rs:445): match a, mismatch -b, N -1."""
    m = np.full((5, 5), -1, dtype=np.int8)
    for i in range(4):
        for j in range(4):
            m[i, j] = a if i == j else -b
    return m.reshape(-1)


# ---- serving -------------------------------------------------------------------------------------
CHUNK_EXTEND, CHUNK_RESCUE = 0, 1
TARGET_FROM_REF, TARGET_REVERSE = 1 << 62, 1 << 63


class Chunk(C.Structure):
    _fields_ = [("kind", C.c_int32), ("ext", C.c_void_p), ("ksw", C.c_void_p), ("n", C.c_size_t),
                ("arena", C.c_void_p), ("arena_len", C.c_size_t), ("q_off", C.c_void_p),
                ("q_len", C.c_void_p), ("t_off", C.c_void_p), ("t_len", C.c_void_p),
                ("h0", C.c_void_p), ("out", C.c_void_p)]


class ServerCounters(C.Structure):
    _fields_ = [("chunks", C.c_uint64), ("jobs", C.c_uint64), ("epochs", C.c_uint64),
                ("peak_device_bytes", C.c_uint64), ("peak_pool_bytes", C.c_uint64),
                ("seconds", C.c_double), ("p50_ms", C.c_double),
                ("p99_ms", C.c_double), ("mean_ms", C.c_double), ("max_ms", C.c_double)]


def _server_lib():
    L = _lib()
    if not hasattr(L, "_server_bound"):
        L.bwg_server_create.restype = C.c_int
        L.bwg_server_create.argtypes = [_P, C.c_uint64, C.c_int, _P]
        L.bwg_server_set_reference.restype = C.c_int
        L.bwg_server_set_reference.argtypes = [_P, _P, C.c_uint64]
        L.bwg_server_submit.restype = C.c_int
        L.bwg_server_submit.argtypes = [_P, _P, _P]
        L.bwg_server_poll.restype = C.c_int
        L.bwg_server_poll.argtypes = [_P, C.c_uint64]
        L.bwg_server_wait.restype = C.c_int
        L.bwg_server_wait.argtypes = [_P, C.c_uint64]
        L.bwg_server_stats.restype = C.c_int
        L.bwg_server_stats.argtypes = [_P, _P, C.c_int]
        L.bwg_server_destroy.restype = None
        L.bwg_server_destroy.argtypes = [_P]
        L._server_bound = True
    return L


class Server:
    """Continuous-batching server on one context (bwg_server_* in include/bwg.h)."""

    def __init__(self, ctx, hbm_budget_bytes=64 << 30, slots=3):
        L = _server_lib()
        h = C.c_void_p()
        _check(L.bwg_server_create(ctx.h, int(hbm_budget_bytes), int(slots), C.byref(h)), "bwg_server_create")
        self.h, self.ctx = h, ctx
        self._live = {}

    def set_reference(self, ref):
        ref = np.ascontiguousarray(ref, dtype=np.uint8)
        _check(_server_lib().bwg_server_set_reference(self.h, ref.ctypes.data, ref.size), "bwg_server_set_reference")

    def submit(self, batch, params, t_off=None):
        """Queue a JobBatch; t_off overrides the batch's target offsets (e.g. genome windows with
        TARGET_FROM_REF).  Returns a ticket; the arrays stay referenced until wait()."""
        is_ext = isinstance(params, ExtParams)
        out = np.zeros(len(batch), dtype=EXT_DTYPE if is_ext else KSW_DTYPE)
        return self.submit_into(batch, params, out, t_off)

    def submit_into(self, batch, params, out, t_off=None):
        """submit() with a caller-owned result array (kept referenced by the caller)."""
        is_ext = isinstance(params, ExtParams)
        t = np.ascontiguousarray(batch.t_off if t_off is None else t_off, dtype=np.uint64)
        ch = Chunk(CHUNK_EXTEND if is_ext else CHUNK_RESCUE,
                   C.addressof(params) if is_ext else None, None if is_ext else C.addressof(params),
                   len(batch), batch.arena.ctypes.data, batch.arena.size, batch.q_off.ctypes.data,
                   batch.q_len.ctypes.data, t.ctypes.data, batch.t_len.ctypes.data,
                   batch.h0.ctypes.data, out.ctypes.data)
        tk = C.c_uint64()
        _check(_server_lib().bwg_server_submit(self.h, C.byref(ch), C.byref(tk)), "bwg_server_submit")
        self._live[tk.value] = (batch, params, t, out)
        return tk.value

    def wait(self, ticket):
        rc = _server_lib().bwg_server_wait(self.h, ticket)
        _, _, _, out = self._live.pop(ticket)
        _check(rc, "bwg_server_wait")
        return out

    def wait_rc(self, ticket):
        """bwg_server_wait's return code (the inputs stay referenced by this object)."""
        return _server_lib().bwg_server_wait(self.h, ticket)

    def poll_rc(self, ticket):
        """bwg_server_poll's return code."""
        return _server_lib().bwg_server_poll(self.h, ticket)

    def counters(self, reset=False):
        c = ServerCounters()
        _check(_server_lib().bwg_server_stats(self.h, C.byref(c), 1 if reset else 0), "bwg_server_stats")
        return {f: getattr(c, f) for f, _ in ServerCounters._fields_}

    def close(self):
        if self.h:
            _server_lib().bwg_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def genome(seed=42):
    """The genome every generated job of `seed` is drawn from (codes 0..4)."""
    W = _wlib()
    W.bwgw_genome_len.restype = C.c_uint64
    W.bwgw_genome.restype = None
    W.bwgw_genome.argtypes = [C.c_uint64, C.c_void_p]
    g = np.zeros(int(W.bwgw_genome_len()), dtype=np.uint8)
    W.bwgw_genome(seed, g.ctypes.data)
    return g


def generate_windows(kind, n, seed=42, genome_seed=None):
    """Like generate(), drawing jobs with `seed` from the genome of `genome_seed`; also returns each
    job's target as a genome window (TARGET_FROM_REF | start, plus TARGET_REVERSE for left
    extensions)."""
    W = _wlib()
    W.bwgw_generate_on.restype = C.c_void_p
    W.bwgw_generate_on.argtypes = [C.c_int, C.c_uint64, C.c_uint64, C.c_uint64]
    W.bwgw_copy_windows.restype = None
    W.bwgw_copy_windows.argtypes = [C.c_void_p, C.c_void_p]
    h = W.bwgw_generate_on(kind, seed if genome_seed is None else genome_seed, seed, n)
    try:
        n = int(W.bwgw_count(h))
        arena = np.zeros(int(W.bwgw_arena_len(h)) + 1, dtype=np.uint8)
        q_off = np.zeros(n, dtype=np.uint64)
        q_len = np.zeros(n, dtype=np.uint32)
        t_off = np.zeros(n, dtype=np.uint64)
        t_len = np.zeros(n, dtype=np.uint32)
        h0 = np.zeros(n, dtype=np.int32)
        kinds = np.zeros(n, dtype=np.uint8)
        win = np.zeros(n, dtype=np.uint64)
        W.bwgw_copy(h, arena.ctypes.data, q_off.ctypes.data, q_len.ctypes.data, t_off.ctypes.data,
                    t_len.ctypes.data, h0.ctypes.data, kinds.ctypes.data)
        W.bwgw_copy_windows(h, win.ctypes.data)
    finally:
        W.bwgw_free(h)
    return JobBatch(arena, q_off, q_len, t_off, t_len, h0, kinds), win


def serve_run(ctx, server, rate_jobs_s, producers=8, duration_s=5.0, seed=42, use_ref=True,
              pool_chunks=64, verify=True):
    """Poisson load from `producers` threads against `server` (the C++ driver in the workload
    library).  Returns p50/p99/mean/max chunk latency (ms), completed jobs/s and counts."""
    W = _wlib()
    W.bwgw_serve_run.restype = C.c_int
    W.bwgw_serve_run.argtypes = [_P, _P, C.c_double, C.c_int, C.c_double, C.c_uint64, C.c_int, C.c_int,
                                 C.c_int, _P]
    out = (C.c_double * 14)()
    _check(W.bwgw_serve_run(ctx.h, server.h, float(rate_jobs_s), int(producers), float(duration_s), seed,
                            1 if use_ref else 0, int(pool_chunks), 1 if verify else 0, out), "bwgw_serve_run")
    keys = ("p50_ms", "p99_ms", "mean_ms", "max_ms", "jobs_per_s", "chunks", "jobs", "epochs",
            "mismatched_chunks", "failed_chunks", "peak_device_bytes", "peak_pool_bytes",
            "ext_jobs_submitted", "rescue_jobs_submitted")
    return dict(zip(keys, list(out)))
