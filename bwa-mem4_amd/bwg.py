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
WORKLOAD_LIB_PATH = os.path.join(HERE, "libbwg_workload.so")

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
OPT_KSW_SCORE_WIDTH = 10
STAT_LAUNCHES, STAT_JOBS, STAT_LAST_KERNEL_NS, STAT_LAST_MAIN_NS, STAT_LAST_MAIN_LAUNCHES = 100, 101, 102, 103, 104

# Every symbol include/bwg.h declares.
EXPORTED_SYMBOLS = (
    "bwg_create", "bwg_destroy", "bwg_name", "bwg_strerror", "bwg_extend_batch", "bwg_extend",
    "bwg_ksw_align2_batch", "bwg_set_option", "bwg_get_stat", "bwg_batch_stage_ext",
    "bwg_batch_stage_ksw", "bwg_batch_run", "bwg_batch_results", "bwg_batch_destroy",
    "bwg_server_create", "bwg_server_set_reference", "bwg_server_submit", "bwg_server_poll",
    "bwg_server_wait", "bwg_server_stats", "bwg_server_destroy", "bwg_hbm_copy", "bwg_tune_report",
    "bwg_batch_ungapped_hits", "bwg_extend_side_batch", "bwg_batch_stage_side", "bwg_batch_side_bands",
)


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

    def side_bands(self):
        """Accepted band of every job of a side batch's last run."""
        w = np.zeros(self.n, dtype=np.int32)
        _check(_lib().bwg_batch_side_bands(self.ctx.h, self.h, w.ctypes.data), "bwg_batch_side_bands")
        return w

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
                ("peak_device_bytes", C.c_uint64), ("seconds", C.c_double), ("p50_ms", C.c_double),
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
    out = (C.c_double * 11)()
    _check(W.bwgw_serve_run(ctx.h, server.h, float(rate_jobs_s), int(producers), float(duration_s), seed,
                            1 if use_ref else 0, int(pool_chunks), 1 if verify else 0, out), "bwgw_serve_run")
    keys = ("p50_ms", "p99_ms", "mean_ms", "max_ms", "jobs_per_s", "chunks", "jobs", "epochs",
            "mismatched_chunks", "failed_chunks", "peak_device_bytes")
    return dict(zip(keys, list(out)))
