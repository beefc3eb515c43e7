"""CPU-side checks of the C ABI library and the workload generators (no GPU compute)."""
import ctypes
import os
import re

import numpy as np

import bwg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    src = open(os.path.join(ROOT, "include", "bwg.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(bwg_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_binding_symbols():
    assert _declared_symbols() == sorted(bwg.EXPORTED_SYMBOLS)


def test_library_loads_and_exports_every_declared_symbol():
    lib = ctypes.CDLL(bwg.LIB_PATH)
    for sym in _declared_symbols():
        assert hasattr(lib, sym), sym


def test_strerror_codes():
    L = bwg.library()
    for code in (0, -5, -7, -11, -12, -19, -22):
        assert L.bwg_strerror(code) and L.bwg_strerror(code) != b"unknown error"
    assert L.bwg_strerror(-1234) == b"unknown error"


def test_struct_layouts_match_header():
    assert ctypes.sizeof(bwg.ExtParams) == 4 + 25 + 3 + 7 * 4
    assert ctypes.sizeof(bwg.KswParams) == 4 + 25 + 3 + 6 * 4
    assert bwg.EXT_DTYPE.itemsize == 24 and bwg.KSW_DTYPE.itemsize == 28


def test_null_arguments_rejected_without_device():
    L = bwg.library()
    assert L.bwg_create(0, None) == bwg.BWG_EINVAL
    assert L.bwg_set_option(None, 1, 0) == bwg.BWG_EINVAL
    assert L.bwg_get_stat(None, 100) == -1


def test_generators_are_deterministic_and_shaped():
    a = bwg.generate(bwg.WORKLOAD_EXT_SMALL, 3000, seed=42)
    b = bwg.generate(bwg.WORKLOAD_EXT_SMALL, 3000, seed=42)
    assert np.array_equal(a.arena, b.arena) and np.array_equal(a.h0, b.h0)
    assert a.q_len.min() >= 1 and a.q_len.max() <= 130
    assert np.array_equal(a.t_len, a.q_len + 100)
    assert a.h0.min() >= 19 and a.h0.max() <= 150
    assert a.arena.max() <= 4
    r = bwg.generate(bwg.WORKLOAD_RESCUE, 3000, seed=42)
    assert np.all(r.q_len == 150) and r.t_len.min() >= 600 and r.t_len.max() <= 1200
    t = bwg.generate(bwg.WORKLOAD_EXT_TAIL, 20000, seed=42)
    frac_short = np.mean(t.q_len < 30)
    frac_long = np.mean(t.q_len >= 150)
    assert 0.66 < frac_short < 0.74 and 0.03 < frac_long < 0.09
    assert t.q_len.max() <= 250 and t.h0.max() <= 250
    m = bwg.generate(bwg.WORKLOAD_SERVE_MIX, 20000, seed=42)
    assert 0.17 < np.mean(m.kind == 1) < 0.23
