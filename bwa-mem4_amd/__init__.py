"""MI355X Smith-Waterman backend for bwa-mem4's extension and mate-rescue path.

The directory name is not a Python identifier; load it by path (the tests, bench.py and
__graft_entry__ put this directory on sys.path and ``import bwg``).
"""
from .bwg import *  # noqa: F401,F403
